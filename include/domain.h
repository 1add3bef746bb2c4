#if !defined(DOMAIN_H)
#define DOMAIN_H

#include <stddef.h>
#include "sbfs_config.h"

/*
 * Domain description.  Same member order and types as the reference's
 * domain_t (include/domain.h:20-30), so a driver compiled against the
 * reference headers links against this library unchanged.
 *
 * `info` is the reference's MPI decomposition handle (SimpleDecomp).  One
 * GPU owns the whole domain, so it is kept as an opaque pointer that this
 * library leaves NULL.
 *
 * The metric arrays live in host memory, exactly as in the reference:
 *   xf   [Nx+1]  cell-face x positions         (XF(i)   = xf[i-1],  i=1..Nx+1)
 *   xc   [Nx+2]  cell-centre x incl. walls     (XC(i)   = xc[i],    i=0..Nx+1)
 *   hxxf [Nx+1]  face scale factor xc(i)-xc(i-1)
 *   hxxc [Nx]    centre scale factor xf(i+1)-xf(i)
 *   jdxf [Nx+1]  hxxf * hy
 *   jdxc [Nx]    hxxc * hy
 * The library mirrors them to device memory on first use.
 */
typedef struct sdecomp_info_t_ sdecomp_info_t;

typedef struct {
  sdecomp_info_t * info;
  size_t glsizes[NDIMS];
  size_t mysizes[NDIMS];
  size_t offsets[NDIMS];
  double lengths[NDIMS];
  double * SBFS_RESTRICT xf, * SBFS_RESTRICT xc;
  double * SBFS_RESTRICT jdxf, * SBFS_RESTRICT jdxc;
  double * SBFS_RESTRICT hxxf, * SBFS_RESTRICT hxxc;
  double hy;
} domain_t;

SBFS_EXTERN_C_BEGIN

/* load glsizes, lengths, xf, xc from an NPY directory and build the metrics
 * (reference: src/domain.c:173-222) */
extern int domain_init(
    const char dirname_ic[],
    domain_t * domain
);

/* write glsizes, lengths, xf, xc as NPY files (reference: src/domain.c:54-75) */
extern int domain_save(
    const char dirname[],
    const domain_t * domain
);

SBFS_EXTERN_C_END

#endif /* DOMAIN_H */
