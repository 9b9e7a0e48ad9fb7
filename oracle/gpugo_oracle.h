/* gpugo_oracle.h -- TEST INFRASTRUCTURE ONLY (see gpugo_oracle.c). */
#ifndef GPUGO_ORACLE_H
#define GPUGO_ORACLE_H
#include <stdint.h>
#include "../include/gpugo.h"

/* oracle_problem.flags */
#define ORACLE_MEAN_VALUE 1   /* also bound by the mean-value form */

typedef struct {
  int32_t d, family, wid, cycle, k_select, n_samples, threads, flags;
  const double *lo, *hi, *params;
  double tol_f, tol_x, init_ub;
  int64_t max_iters;
} oracle_problem;

int oracle_rounding_selftest(void);
int oracle_ia_eval(int op, int64_t n, const double *alo, const double *ahi,
                   const double *blo, const double *bhi, double s,
                   double *olo, double *ohi);
void oracle_decode(int d, const double *plo, const double *phi, int win, int wid,
                   int c, int64_t n, const uint32_t *pos, double *clo, double *chi);
int oracle_eval_boxes(int fam, int d, const double *P, int64_t n,
                      const double *lo, const double *hi,
                      const double *dlo, const double *dhi, int S,
                      double *flo, double *fhi, double *glo, double *ghi,
                      int32_t *kkt, double *ub);
int oracle_mv_lower(int fam, int d, const double *P, int64_t n,
                    const double *lo, const double *hi, double *out);
int oracle_solve(const oracle_problem *pr, gpugo_result *res,
                 double *out_lo, double *out_hi, double *out_lb, int64_t out_max);
int64_t oracle_split_bound(int fam, int d, const double *P, int64_t nparents,
                           const double *plo, const double *phi, const int32_t *pwin,
                           int wid, int cycle, int S, const double *dlo, const double *dhi,
                           double *ub_inout, int64_t out_max, uint32_t *out_parent,
                           uint32_t *out_pos, double *out_lb);
#endif
