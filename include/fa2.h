/*
 * fa2.h — the drop-in C ABI of the ForceAtlas2 GPU backend, implemented for
 * AMD Instinct MI355X (gfx950) by libfa2hip.so.
 *
 * Every declaration below keeps the name, the argument list, the types and
 * the meaning of the reference header src/fa2.h of forceatlas2r 0.3.0, so
 * that the package's C code (src/forceatlas2.c) links against this library
 * in place of src/gpu.c + src/gpu_full.c + src/ocl_loader.c without change.
 * Each entry point cites the reference declaration it replaces.
 *
 * Differences of behaviour, all on the side of the implementation:
 *  - the device is a HIP device (hipGetDeviceCount), not an OpenCL one;
 *  - fa2_gpu_error() is thread-local, so concurrent sessions are safe;
 *  - fa2_full_run() ignores `source` (kernels are compiled ahead of time) and
 *    keeps the adaptive speed on the device: there is no host
 *    synchronisation inside the iteration loop.
 */
#ifndef FA2_H
#define FA2_H

#ifdef __cplusplus
extern "C" {
#endif

/* Barnes-Hut region of the host-built Gephi quadtree (src/fa2.h:21-27).
 * A region covers perm[start .. start+count-1]; its children are
 * child .. child+nchild-1; `skip` is the next region of a stackless
 * traversal that is not a descendant (-1 at the end). */
typedef struct {
    double mass, cx, cy, size;
    int start, count;
    int child, nchild;
    int skip;
    int buf;
} fa2_region;

/* A node while the tree is built (src/fa2.h:30-33). */
typedef struct {
    double x, y, m;
    int idx;
} fa2_node;

/* The tree (src/fa2.h:35-44). px/py/pm/ps hold the nodes in tree order. */
typedef struct {
    fa2_region *reg;
    int nreg;
    int *perm;
    int *cnt;
    fa2_node *buf[2];
    double *px, *py, *pm, *ps;
} fa2_tree;

/* Device description (src/fa2.h:50-58). */
typedef struct {
    char name[128];
    char vendor[128];
    char platform[128];
    char version[128];
    int is_gpu;
    int has_double;
    double memory; /* bytes */
} fa2_gpu_devinfo;

typedef struct fa2_gpu fa2_gpu;

/* Last error of the calling thread; never NULL (src/fa2.h:63). */
const char *fa2_gpu_error(void);

/* Number of HIP devices, 0 when there is no usable runtime (src/fa2.h:66). */
int fa2_gpu_ndevices(void);

/* Description of device idx (0-based); 0 on success (src/fa2.h:68). */
int fa2_gpu_device_info(int idx, fa2_gpu_devinfo *info);

/* Open a repulsion session for n nodes, 1 <= n <= 2^27 (src/fa2.h:72-73).
 * NULL on error, with the reason in fa2_gpu_error(). `threads` is the number
 * of host threads used to convert the data. */
fa2_gpu *fa2_gpu_open(int device, int use_double, int n, int adjust,
                      int barnes_hut, int threads);

/* Seconds spent sending, computing and fetching (src/fa2.h:76). */
void fa2_gpu_times(const fa2_gpu *g, double times[3]);

/* Release a session; NULL is accepted (src/fa2.h:77). */
void fa2_gpu_close(fa2_gpu *g);

/* Exact repulsion between every pair of nodes, ADDED to dx and dy; size may
 * be NULL when node sizes are not used. 0 on success, 1 on error
 * (src/fa2.h:81-83). */
int fa2_gpu_repulse_exact(fa2_gpu *g, const double *x, const double *y,
                          const double *mass, const double *size,
                          double scaling, double *dx, double *dy);

/* Barnes-Hut repulsion over the host-built tree t, ADDED to dx[perm[k]] and
 * dy[perm[k]]. 0 on success, 1 on error (src/fa2.h:84-85). */
int fa2_gpu_repulse_bh(fa2_gpu *g, const fa2_tree *t, double scaling,
                       double theta, double *dx, double *dy);

/* Release everything cached between calls (src/fa2.h:88). */
void fa2_gpu_shutdown(void);

/* Monotonic clock in seconds (src/fa2.h:91). */
double fa2_now(void);

/* Arguments of the all-GPU layout (src/fa2.h:97-115). */
typedef struct {
    int n;
    const int *inc_off, *inc_node;
    const double *inc_coef;
    const double *mass;
    const double *size;
    const int *fixed;
    int any_fixed;
    double scaling, gravity, theta, jitter_tolerance;
    int strong, linlog, adjust;
    int iterations;
    int device, use_double;
    int profile;
    const char *source;
    int (*interrupted)(void);
} fa2_full_params;

/* Slots of the `times` argument of fa2_full_run (src/fa2.h:117-126). */
enum {
    FA2_FULL_T_TOTAL,
    FA2_FULL_T_CODES,
    FA2_FULL_T_SORT,
    FA2_FULL_T_TREE,
    FA2_FULL_T_REPULSION,
    FA2_FULL_T_ATTRACTION,
    FA2_FULL_T_MOVE,
    FA2_FULL_NTIMES
};

/* All-GPU layout of n nodes, 2 <= n <= 2^26 (src/fa2.h:131). pos holds x
 * then y and receives the result; times receives FA2_FULL_NTIMES values in
 * seconds. 0 on success, 1 on error. */
int fa2_full_run(const fa2_full_params *p, double *pos, double *times);

/* Adaptive global speed (src/fa2.h:135-137, defined by the reference in
 * src/forceatlas2.c:445-485). libfa2hip.so exports a weak definition so that
 * the host driver of fa2_hip.h works on its own; the package's own
 * definition takes precedence when both are linked. */
void fa2_update_speed(int n, double total_swinging, double total_traction,
                      double jitter_tolerance, double *speed,
                      double *speed_efficiency);

#ifdef __cplusplus
}
#endif

#endif
