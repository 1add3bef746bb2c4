#if !defined(SBFS_CONFIG_H)
#define SBFS_CONFIG_H

/*
 * Common compile-time settings shared by every public header.
 *
 * The solver is two-dimensional: x (index i) is wall-bounded and is the
 * gravity direction, y (index j) is periodic.  The reference selects this
 * with -DNDIMS=2 (Makefile:2); here it is fixed.
 */
#if !defined(NDIMS)
#define NDIMS 2
#endif

#if NDIMS != 2
#error "only the two-dimensional solver is provided"
#endif

/* `restrict` is a C99 keyword; C++ spells it `__restrict`. */
#if defined(__cplusplus)
#define SBFS_RESTRICT __restrict
#define SBFS_EXTERN_C_BEGIN extern "C" {
#define SBFS_EXTERN_C_END }
#else
#define SBFS_RESTRICT restrict
#define SBFS_EXTERN_C_BEGIN
#define SBFS_EXTERN_C_END
#endif

#endif /* SBFS_CONFIG_H */
