/*
 * ORACLE — test infrastructure only. Declarations of the CPU restatement of
 * the reference ForceAtlas2 algorithms (fa2_oracle.c). Loaded by tests/,
 * the smoke check and the CPU baseline of bench.py, never by the product.
 */
#ifndef FA2_ORACLE_H
#define FA2_ORACLE_H
#include <stdint.h>

/* Arguments of the all-GPU layout, as in fa2_full_params (src/fa2.h:97-115)
 * without the device fields. */
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
} orc_large_params;

double orc_now(void);
int orc_has_openmp(void);
void orc_update_speed(int n, double sw, double tr, double jt_user,
                      double *speed, double *eff);
int orc_gephi_tree(int n, const double *x, const double *y, const double *mass,
                   const double *size, int *nreg, double *reg_d, int *reg_i,
                   int *perm, double *px, double *py, double *pm, double *ps);
int orc_layout(int n, int m, const int *from, const int *to, const double *w,
               double *pos, const double *size, const int *fixed,
               int iterations, const double *dpar, const int *flags,
               int threads, double *trace_pos, double *trace_force,
               double *times);
void orc_repulse_exact(int n, const double *x, const double *y,
                       const double *mass, const double *size, double scaling,
                       int threads, double *dx, double *dy);
int orc_repulse_bh(int n, const double *x, const double *y, const double *mass,
                   const double *size, double scaling, double theta,
                   int threads, double *dx, double *dy);
/* FP64 model of the single-precision offload: float32 inputs and records,
 * float32 acceptance test, FP64 terms (fa2_oracle.c). */
int orc_repulse_bh_single(int n, const double *x, const double *y, const double *mass,
                          const double *size, double scaling, double theta, double *dx,
                          double *dy);
uint32_t orc_spread16(uint32_t v);
void orc_radix_sort(int n, const uint32_t *key_in, uint32_t *key_out,
                    uint32_t *perm);
void orc_karras(int n, const uint32_t *key, int *child, int *range,
                int *parent);
int orc_large_prepare(int n, int m, const int *from, const int *to,
                      const double *w, int outbound, double *mass,
                      int *inc_off, int *inc_node, double *inc_coef);

#define ORC_DECLARE_LARGE(S, PR)                                              \
    void orc_morton_##S(int n, const PR *x, const PR *y, uint32_t *key);      \
    void orc_large_tree_##S(int n, const PR *x, const PR *y, uint32_t *key,   \
                            uint32_t *key_sorted, uint32_t *order,            \
                            int *child, int *range, int *parent);             \
    int orc_large_regions_##S(int n, const PR *x, const PR *y,                \
                              const double *mass, double *ni);                \
    int orc_large_run_##S(const orc_large_params *p, double *pos,             \
                          int iterations, double *trace_pos,                  \
                          double *trace_force, double *trace_speed,           \
                          double *trace_rep);
ORC_DECLARE_LARGE(f, float)
ORC_DECLARE_LARGE(d, double)
ORC_DECLARE_LARGE(m, double)

#endif
