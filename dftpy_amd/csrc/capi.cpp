// C ABI of libofdft_hip.so: plan lifetime, FFTs with DFTpy's normalisation,
// functional settings and evaluation, reductions, local PP / Ewald setup.
// Every entry point catches exceptions and reports them through
// ofdft_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "internal.h"

namespace ofdft {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

std::recursive_mutex &legacy_lock() {
    static std::recursive_mutex m;
    return m;
}

void local_pp(Plan &P, int nspecies, const double *q0, const double *dq, const int *ncoef,
              const double *const *coef, const double *qmax, const double *vsmall, const double *dqs, int nat,
              const double *pos_cart, const double *pos_frac, const int *species, int pme, int order, void *vreal,
              hipStream_t s, const double *screen_z = nullptr, const double *screen_r1 = nullptr,
              const double *screen_r2 = nullptr, const double *resp = nullptr);
void tn_precond_fill(Plan &P, int n, double dq, double rho0, const double *resp, hipStream_t s);
void pp_forces(Plan &P, const double *rho, int nspecies, const double *q0, const double *dq, const int *ncoef,
               const double *const *coef, const double *qmax, const double *vsmall, const double *dqs, int nat,
               const double *pos_cart, const double *pos_frac, const int *species, double *F, hipStream_t s);
void pp_stress(Plan &P, const double *rho, int nspecies, const double *q0, const double *dq, const int *ncoef,
               const double *const *coef, const double *qmax, const double *vsmall, const double *dqs, int nat,
               const double *pos_cart, const double *pos_frac, const int *species, double energy, double *S,
               hipStream_t s);
void ewald_fs(Plan &P, int nat, const double *pos_cart, const double *pos_frac, const double *charges,
              double precision, double *F, double *S, hipStream_t s);
void pp_forces_pme(Plan &P, const double *rho, int nspecies, const double *q0, const double *dq, const int *ncoef,
                   const double *const *coef, const double *qmax, const double *vsmall, const double *dqs, int nat,
                   const double *pos_cart, const double *pos_frac, const int *species, int order, double *F,
                   hipStream_t s);
void pp_stress_pme(Plan &P, const double *rho, int nspecies, const double *q0, const double *dq, const int *ncoef,
                   const double *const *coef, const double *qmax, const double *vsmall, const double *dqs, int nat,
                   const double *pos_cart, const double *pos_frac, const int *species, int order, double energy,
                   double *S, hipStream_t s);
void ewald_fs_pme(Plan &P, int nat, const double *pos_cart, const double *pos_frac, const double *charges,
                  double precision, int order, double *F, double *S, hipStream_t s);
void functional_stress(Plan &P, const double *rho, const double *energies, double *out, hipStream_t s);
void ewald(Plan &P, int nat, const double *pos_cart, const double *pos_frac, const double *charges, int pme,
           int order, double precision, double *out, hipStream_t s);

template <typename F> static int guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("unknown error");
        return 1;
    }
}

static void inv3(const double *a, double *r) {
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                       a[2] * (a[3] * a[7] - a[4] * a[6]);
    OFDFT_REQUIRE(std::fabs(det) > 0, "singular lattice");
    r[0] = (a[4] * a[8] - a[5] * a[7]) / det;
    r[1] = (a[2] * a[7] - a[1] * a[8]) / det;
    r[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    r[3] = (a[5] * a[6] - a[3] * a[8]) / det;
    r[4] = (a[0] * a[8] - a[2] * a[6]) / det;
    r[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    r[6] = (a[3] * a[7] - a[4] * a[6]) / det;
    r[7] = (a[1] * a[6] - a[0] * a[7]) / det;
    r[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

// lattice rows a_i (bohr) -> B = 2 pi inv(A)^T (grid.py:405-414), volume, dV
static void set_lattice(Plan &P, const double *lattice) {
    std::memcpy(P.lat, lattice, sizeof(P.lat));
    double inv[9];
    inv3(lattice, inv);
    // row i of B is column i of inv(A)
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P.B[3 * i + j] = 2.0 * M_PI * inv[3 * j + i];
    const double *a = lattice;
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                       a[2] * (a[3] * a[7] - a[4] * a[6]);
    P.volume = std::fabs(det);
    P.dV = P.volume / (double)P.N;
}

static hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

void plan_aux_setup(Plan &P);
void plan_aux_free(Plan &P);

}  // namespace ofdft

using namespace ofdft;

struct ofdft_plan : public Plan {};

extern "C" {

int ofdft_version(void) { return 1; }

const char *ofdft_last_error(void) { return g_err.c_str(); }

int ofdft_plan_create(const int nr[3], const double lattice[9], int precision, int full, int max_batch,
                      ofdft_plan **out) {
    return guard([&] {
        OFDFT_LEGACY_LOCK;
        OFDFT_REQUIRE(out != nullptr, "plan_create: null output pointer");
        OFDFT_REQUIRE(precision == OFDFT_FP64 || precision == OFDFT_FP32, "precision must be FP64 or FP32");
        for (int d = 0; d < 3; ++d) OFDFT_REQUIRE(nr[d] >= 1 && nr[d] <= 4096, "grid size out of range");
        ofdft_plan *P = new ofdft_plan();
        try {
            std::memset(P->tw, 0, sizeof(P->tw));
            P->spec = P->work = P->aux = nullptr;
            P->red = P->scal = nullptr;
            P->pinned = P->pinned_dev = nullptr;
            for (int d = 0; d < 3; ++d) P->nr[d] = nr[d];
            P->nzc = nr[2] / 2 + 1;
            {
                const int per = precision == OFDFT_FP64 ? 8 : 16;  // complex elements per 128 bytes
                P->nzp = ((P->nzc + per - 1) / per) * per;
            }
            P->N = (long long)nr[0] * nr[1] * nr[2];
            P->Nc = (long long)nr[0] * nr[1] * P->nzc;
            P->Nfull = P->N;
            set_lattice(*P, lattice);
            P->prec = precision;
            P->full = full ? 1 : 0;
            P->max_batch = max_batch < 1 ? 1 : max_batch;
            plan_fft_setup(*P);
            const size_t rs = precision == OFDFT_FP64 ? 8 : 4;
            const int nslot = P->max_batch > 4 ? P->max_batch : 4;
            // spectra: nslot half grids, or nslot/2 full complex grids
            const size_t ncp = (size_t)nr[0] * nr[1] * P->nzp;
            const size_t spec_bytes = std::max((size_t)nslot * ncp, (size_t)(2 * P->N)) * 2 * rs;
            OFDFT_HIP(hipMalloc(&P->spec, spec_bytes));
            // work: 4 real fields, at least one FP64 field for setup arrays
            const size_t work_bytes = std::max((size_t)(4 * P->N * rs), (size_t)(P->N * 8));
            OFDFT_HIP(hipMalloc(&P->work, work_bytes));
            P->red_blocks = 4096;
            OFDFT_HIP(hipMalloc(&P->red, sizeof(double) * 4096 * 16));
            OFDFT_HIP(hipMalloc(&P->scal, sizeof(double) * kScalSlots));
            OFDFT_HIP(hipMemset(P->scal, 0, sizeof(double) * kScalSlots));
            OFDFT_HIP(hipHostMalloc((void **)&P->pinned, 128 * sizeof(double), hipHostMallocMapped));
            OFDFT_HIP(hipHostGetDevicePointer((void **)&P->pinned_dev, P->pinned, 0));
            P->ws_bytes = spec_bytes + work_bytes + sizeof(double) * (4096 * 16 + kScalSlots);
            plan_aux_setup(*P);
            P->ws_bytes += P->aux_bytes;
            OFDFT_HIP(hipDeviceSynchronize());
        } catch (...) {
            ofdft_plan_destroy(P);
            throw;
        }
        *out = P;
    });
}

int ofdft_plan_autotune(ofdft_plan *plan, int *variants, float *times, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(plan != nullptr, "plan_autotune: null plan");
        plan_autotune(*plan, S(stream), times);
        if (variants) {
            variants[0] = plan->var_r2c;
            variants[1] = plan->var_c2r;
            variants[2] = plan->var_cols[0];
            variants[3] = plan->var_cols[1];
        }
    });
}

int ofdft_plan_set_variants(ofdft_plan *plan, const int *variants) {
    return guard([&] {
        OFDFT_REQUIRE(plan != nullptr && variants != nullptr, "plan_set_variants: null pointer");
        for (int i = 0; i < 4; ++i)
            OFDFT_REQUIRE(variants[i] >= 0 && variants[i] < (i < 2 ? kTileVariants : kColVariants),
                          "variant out of range");
        plan->var_r2c = variants[0];
        plan->var_c2r = variants[1];
        plan->var_cols[0] = variants[2];
        plan->var_cols[1] = variants[3];
    });
}

int ofdft_plan_set_lattice(ofdft_plan *plan, const double lattice[9]) {
    return guard([&] {
        OFDFT_REQUIRE(plan != nullptr && lattice != nullptr, "plan_set_lattice: null pointer");
        set_lattice(*plan, lattice);
    });
}

int ofdft_plan_destroy(ofdft_plan *P) {
    if (!P) return 0;
    return guard([&] {
        OFDFT_LEGACY_LOCK;
        plan_fft_free(*P);
        plan_aux_free(*P);
        lmgp_free(*P);
        if (P->spec) (void)hipFree(P->spec);
        if (P->work) (void)hipFree(P->work);
        if (P->red) (void)hipFree(P->red);
        if (P->scal) (void)hipFree(P->scal);
        if (P->pinned) (void)hipHostFree(P->pinned);
        if (P->tn_pre) (void)hipFree(P->tn_pre);
        if (P->lbfgs_buf) (void)hipFree(P->lbfgs_buf);
        plan_setup_free(*P);
        delete P;
    });
}

size_t ofdft_plan_workspace_bytes(const ofdft_plan *P) { return P ? P->ws_bytes : 0; }

int ofdft_set_tn_precond(ofdft_plan *P, int n, double dq, const double *table) {
    return guard([&] {
        OFDFT_LEGACY_LOCK;
        OFDFT_REQUIRE(P != nullptr, "set_tn_precond: null plan");
        if (n <= 0) {
            P->tn_pre_n = 0;
            return;
        }
        OFDFT_REQUIRE(n >= 2 && dq > 0.0 && table != nullptr, "set_tn_precond: bad table");
        if (n > P->tn_pre_cap) {
            if (P->tn_pre) OFDFT_HIP(hipFree(P->tn_pre));
            P->tn_pre = nullptr;
            P->tn_pre_cap = 0;
            OFDFT_HIP(hipMalloc(&P->tn_pre, sizeof(double) * n));
            P->tn_pre_cap = n;
        }
        OFDFT_HIP(hipMemcpy(P->tn_pre, table, sizeof(double) * n, hipMemcpyHostToDevice));
        P->tn_pre_n = n;
        P->tn_pre_dq = dq;
    });
}

int ofdft_plan_reserve_lbfgs(ofdft_plan *P, int maxcor) {
    return guard([&] {
        OFDFT_LEGACY_LOCK;
        OFDFT_REQUIRE(P != nullptr, "plan_reserve_lbfgs: null plan");
        OFDFT_REQUIRE(maxcor >= 0 && maxcor + 1 <= 16, "plan_reserve_lbfgs: maxcor must be in [0, 15]");
        const int pairs = maxcor + 1;
        if (pairs <= P->lbfgs_pairs) return;
        const size_t rs = P->prec == OFDFT_FP64 ? 8 : 4;
        if (P->lbfgs_buf) OFDFT_HIP(hipFree(P->lbfgs_buf));
        P->lbfgs_buf = nullptr;
        P->lbfgs_pairs = 0;
        OFDFT_HIP(hipMalloc(&P->lbfgs_buf, (size_t)2 * pairs * P->N * rs));
        P->lbfgs_pairs = pairs;
    });
}

int ofdft_fft_r2c(ofdft_plan *P, const void *in, void *out, int batch, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && in && out && batch >= 1, "fft_r2c: bad arguments");
        if (P->prec == OFDFT_FP64)
            fft_r2c<double>(*P, (const double *)in, (double2 *)out, batch, P->N, P->Nc, P->dV, nullptr, nullptr,
                            S(stream));
        else
            fft_r2c<float>(*P, (const float *)in, (float2 *)out, batch, P->N, P->Nc, P->dV, nullptr, nullptr,
                           S(stream));
    });
}

int ofdft_fft_c2r(ofdft_plan *P, const void *in, void *out, int batch, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && in && out && batch >= 1, "fft_c2r: bad arguments");
        const int nslot = P->max_batch > 4 ? P->max_batch : 4;
        OFDFT_REQUIRE(batch <= nslot, "fft_c2r: batch exceeds the plan's max_batch");
        const double sc = 1.0 / ((double)P->N * P->dV);
        if (P->prec == OFDFT_FP64)
            fft_c2r<double>(*P, (const double2 *)in, (double2 *)P->spec, (double *)out, batch, P->Nc, P->Nc, P->N,
                            sc, S(stream));
        else
            fft_c2r<float>(*P, (const float2 *)in, (float2 *)P->spec, (float *)out, batch, P->Nc, P->Nc, P->N, sc,
                           S(stream));
    });
}

int ofdft_fft_c2c(ofdft_plan *P, const void *in, void *out, int inverse, int batch, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && in && out && batch >= 1, "fft_c2c: bad arguments");
        const double sc = inverse ? 1.0 / ((double)P->N * P->dV) : P->dV;
        if (P->prec == OFDFT_FP64)
            fft_c2c<double>(*P, (const double2 *)in, (double2 *)out, inverse != 0, batch, P->N, sc, S(stream));
        else
            fft_c2c<float>(*P, (const float2 *)in, (float2 *)out, inverse != 0, batch, P->N, sc, S(stream));
    });
}

int ofdft_plan_set_fft_kernels(ofdft_plan *P, int fixed) {
    return guard([&] {
        OFDFT_REQUIRE(P, "null plan");
        P->use_fixed = fixed != 0;
    });
}

int ofdft_set_kedf(ofdft_plan *P, double x, double y, double alpha, double beta, double rho0, double sigma) {
    return guard([&] {
        OFDFT_REQUIRE(P, "null plan");
        P->x = x;
        P->y = y;
        P->alpha = alpha;
        P->beta = beta;
        P->rho0 = rho0;
        P->sigma = sigma;
    });
}

int ofdft_set_terms(ofdft_plan *P, unsigned terms) {
    return guard([&] {
        OFDFT_REQUIRE(P, "null plan");
        OFDFT_REQUIRE((terms & ~127u) == 0, "unknown term bits");
        OFDFT_REQUIRE(!((terms & OFDFT_TERM_WT) && (terms & OFDFT_TERM_LMGP)), "WT and LMGP are exclusive");
        P->terms = terms;
    });
}

int ofdft_set_core(ofdft_plan *P, const void *core) {
    return guard([&] {
        OFDFT_REQUIRE(P != nullptr, "set_core: null plan");
        P->core = core;
    });
}

int ofdft_set_vloc(ofdft_plan *P, const void *vreal) {
    return guard([&] {
        OFDFT_REQUIRE(P, "null plan");
        P->vloc = vreal;
    });
}

int ofdft_set_eion(ofdft_plan *P, double eion) {
    return guard([&] {
        OFDFT_REQUIRE(P, "null plan");
        P->eion = eion;
    });
}

int ofdft_eval(ofdft_plan *P, const void *rho, unsigned calc, void *pot, double *energies, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && rho, "eval: bad arguments");
        OFDFT_REQUIRE(!((P->terms & OFDFT_TERM_PP) && P->vloc == nullptr), "eval: PP term set but no vloc");
        if (P->prec == OFDFT_FP64)
            eval<double>(*P, (const double *)rho, calc, (double *)pot, energies, S(stream));
        else
            eval<float>(*P, (const float *)rho, calc, (float *)pot, energies, S(stream));
    });
}

int ofdft_eval_fields(ofdft_plan *P, const void *rho, unsigned calc, void *pot, double *energies, void *edens,
                      void *v2rho2, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && rho && (edens || v2rho2), "eval_fields: bad arguments");
        OFDFT_REQUIRE(!((P->terms & OFDFT_TERM_PP) && P->vloc == nullptr), "eval_fields: PP term set but no vloc");
        if (P->prec == OFDFT_FP64)
            eval_fields<double>(*P, (const double *)rho, calc, (double *)pot, energies, (double *)edens,
                                (double *)v2rho2, S(stream));
        else
            eval_fields<float>(*P, (const float *)rho, calc, (float *)pot, energies, (float *)edens, (float *)v2rho2,
                               S(stream));
    });
}

int ofdft_dot(ofdft_plan *P, const void *a, const void *b, double *out, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && a && out, "dot: bad arguments");
        if (P->prec == OFDFT_FP64)
            dot<double>(*P, (const double *)a, (const double *)b, nullptr, P->N, out, 1.0, S(stream));
        else
            dot<float>(*P, (const float *)a, (const float *)b, nullptr, P->N, out, 1.0, S(stream));
    });
}

int ofdft_bench_pass(ofdft_plan *P, int which, int batch, int iters, const void *rho, double *avg_ms,
                     void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && avg_ms && iters >= 1 && batch >= 1 && batch <= 3, "bench_pass: bad arguments");
        OFDFT_REQUIRE(which >= 0 && which <= 7, "bench_pass: unknown pass");
        OFDFT_REQUIRE((which != 0 && which != 5 && which != 7) || rho != nullptr, "bench_pass: this pass needs rho");
        hipStream_t s = S(stream);
        hipEvent_t e0, e1;
        OFDFT_HIP(hipEventCreate(&e0));
        OFDFT_HIP(hipEventCreate(&e1));
        auto run = [&]() {
            if (P->prec == OFDFT_FP64) {
                double2 *spec = (double2 *)P->spec;
                if (which == 0) {
                    Prologue pro{};
                    pro.code[0] = OP_SQRT_RHOM;
                    pro.code[1] = OP_POW_BETA;
                    pro.beta = P->beta;
                    pro.b56 = std::fabs(P->beta - 5.0 / 6.0) < 1e-14;
                    bench_rows_r2c<double>(*P, (const double *)rho, spec, batch, &pro, s);
                } else if (which == 1 || which == 2) {
                    const long long pt = pipe_pitch(*P);
                    fft_cols<double>(*P, which == 1 ? 1 : 0, false, false, spec, spec, batch,
                                     (long long)P->nr[0] * P->nr[1] * pt, 1.0, s, pt);
                } else if (which == 3) {
                    bench_rows_c2r<double>(*P, spec, (double *)P->work, batch, s);
                } else if (which == 5) {
                    eval<double>(*P, (const double *)rho, OFDFT_CALC_E | OFDFT_CALC_V, (double *)P->aux,
                                 P->scal + 100, s);
                } else if (which == 6 || which == 7) {
                    bench_eval_stage<double>(*P, which - 6, (const double *)rho, (double *)P->aux, s);
                } else {
                    fft_c2c<double>(*P, spec, spec, false, 1, P->N, 1.0, s);
                }
            } else {
                float2 *spec = (float2 *)P->spec;
                if (which == 0) {
                    Prologue pro{};
                    pro.code[0] = OP_SQRT_RHOM;
                    pro.code[1] = OP_POW_BETA;
                    pro.beta = P->beta;
                    pro.b56 = std::fabs(P->beta - 5.0 / 6.0) < 1e-14;
                    bench_rows_r2c<float>(*P, (const float *)rho, spec, batch, &pro, s);
                } else if (which == 1 || which == 2) {
                    const long long pt = pipe_pitch(*P);
                    fft_cols<float>(*P, which == 1 ? 1 : 0, false, false, spec, spec, batch,
                                    (long long)P->nr[0] * P->nr[1] * pt, 1.0, s, pt);
                } else if (which == 3) {
                    bench_rows_c2r<float>(*P, spec, (float *)P->work, batch, s);
                } else if (which == 5) {
                    eval<float>(*P, (const float *)rho, OFDFT_CALC_E | OFDFT_CALC_V, (float *)P->aux, P->scal + 100,
                                s);
                } else if (which == 6 || which == 7) {
                    bench_eval_stage<float>(*P, which - 6, (const float *)rho, (float *)P->aux, s);
                } else {
                    fft_c2c<float>(*P, spec, spec, false, 1, P->N, 1.0, s);
                }
            }
        };
        run();
        OFDFT_HIP(hipEventRecord(e0, s));
        for (int i = 0; i < iters; ++i) run();
        OFDFT_HIP(hipEventRecord(e1, s));
        OFDFT_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        OFDFT_HIP(hipEventElapsedTime(&ms, e0, e1));
        *avg_ms = (double)ms / iters;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int ofdft_local_pp(ofdft_plan *P, int nspecies, const double *q0, const double *dq, const int *ncoef,
                   const double *const *coef, const double *qmax, const double *vsmall, const double *dqs, int nat,
                   const double *pos_cart, const double *pos_frac, const int *species, int pme, int order,
                   void *vreal, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && vreal && nat >= 1, "local_pp: bad arguments");
        for (int a = 0; a < nat; ++a)
            OFDFT_REQUIRE(species[a] >= 0 && species[a] < nspecies, "local_pp: species index out of range");
        local_pp(*P, nspecies, q0, dq, ncoef, coef, qmax, vsmall, dqs, nat, pos_cart, pos_frac, species, pme, order,
                 vreal, S(stream));
    });
}

int ofdft_screened_superposition(ofdft_plan *P, int nspecies, const double *q0, const double *dq, const int *ncoef,
                                 const double *const *coef, const double *qmax, const double *vsmall,
                                 const double *dqs, const double *zval, const double *r1, const double *r2,
                                 const double *resp, int nat, const double *pos_cart, const double *pos_frac,
                                 const int *species, int pme, int order, void *out, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && out && nat >= 1 && zval && r1 && r2 && resp, "screened_superposition: bad arguments");
        for (int a = 0; a < nat; ++a)
            OFDFT_REQUIRE(species[a] >= 0 && species[a] < nspecies, "screened_superposition: species out of range");
        local_pp(*P, nspecies, q0, dq, ncoef, coef, qmax, vsmall, dqs, nat, pos_cart, pos_frac, species, pme, order,
                 out, S(stream), zval, r1, r2, resp);
    });
}

int ofdft_set_tn_precond_uniform(ofdft_plan *P, int n, double dq, double rho0, const double *resp, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P != nullptr && n >= 2 && dq > 0.0 && rho0 > 0.0 && resp && resp[3] > 0.0,
                      "set_tn_precond_uniform: bad arguments");
        if (n > P->tn_pre_cap) {
            OFDFT_LEGACY_LOCK;
            if (P->tn_pre) OFDFT_HIP(hipFree(P->tn_pre));
            P->tn_pre = nullptr;
            P->tn_pre_cap = 0;
            OFDFT_HIP(hipMalloc(&P->tn_pre, sizeof(double) * n));
            P->tn_pre_cap = n;
        }
        tn_precond_fill(*P, n, dq, rho0, resp, S(stream));
        P->tn_pre_n = n;
        P->tn_pre_dq = dq;
    });
}

int ofdft_get_tn_precond(ofdft_plan *P, double *out, int n, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P != nullptr && out != nullptr && n >= 1 && n <= P->tn_pre_n,
                      "get_tn_precond: no table of that length");
        OFDFT_HIP(hipMemcpyAsync(out, P->tn_pre, sizeof(double) * n, hipMemcpyDeviceToDevice, S(stream)));
    });
}

int ofdft_ewald(ofdft_plan *P, int nat, const double *pos_cart, const double *pos_frac, const double *charges,
                int pme, int order, double precision, double *out, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && nat >= 1 && out, "ewald: bad arguments");
        ewald(*P, nat, pos_cart, pos_frac, charges, pme, order, precision, out, S(stream));
    });
}

int ofdft_local_pp_forces(ofdft_plan *P, const void *rho, int nspecies, const double *q0, const double *dq,
                          const int *ncoef, const double *const *coef, const double *qmax, const double *vsmall,
                          const double *dqs, int nat, const double *pos_cart, const double *pos_frac,
                          const int *species, double *forces, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && rho && forces && nat >= 1 && nspecies >= 1, "local_pp_forces: bad arguments");
        for (int a = 0; a < nat; ++a)
            OFDFT_REQUIRE(species[a] >= 0 && species[a] < nspecies, "local_pp_forces: species index out of range");
        for (int k = 0; k < 3 * nat; ++k) forces[k] = 0.0;
        pp_forces(*P, (const double *)rho, nspecies, q0, dq, ncoef, coef, qmax, vsmall, dqs, nat, pos_cart, pos_frac,
                  species, forces, S(stream));
    });
}

int ofdft_local_pp_stress(ofdft_plan *P, const void *rho, int nspecies, const double *q0, const double *dq,
                          const int *ncoef, const double *const *coef, const double *qmax, const double *vsmall,
                          const double *dqs, int nat, const double *pos_cart, const double *pos_frac,
                          const int *species, double energy, double *stress, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && rho && stress && nat >= 1 && nspecies >= 1, "local_pp_stress: bad arguments");
        for (int a = 0; a < nat; ++a)
            OFDFT_REQUIRE(species[a] >= 0 && species[a] < nspecies, "local_pp_stress: species index out of range");
        pp_stress(*P, (const double *)rho, nspecies, q0, dq, ncoef, coef, qmax, vsmall, dqs, nat, pos_cart, pos_frac,
                  species, energy, stress, S(stream));
    });
}

int ofdft_local_pp_forces_pme(ofdft_plan *P, const void *rho, int nspecies, const double *q0, const double *dq,
                              const int *ncoef, const double *const *coef, const double *qmax, const double *vsmall,
                              const double *dqs, int nat, const double *pos_cart, const double *pos_frac,
                              const int *species, int order, double *forces, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && rho && forces && nat >= 1 && nspecies >= 1, "local_pp_forces_pme: bad arguments");
        for (int a = 0; a < nat; ++a)
            OFDFT_REQUIRE(species[a] >= 0 && species[a] < nspecies,
                          "local_pp_forces_pme: species index out of range");
        for (int k = 0; k < 3 * nat; ++k) forces[k] = 0.0;
        pp_forces_pme(*P, (const double *)rho, nspecies, q0, dq, ncoef, coef, qmax, vsmall, dqs, nat, pos_cart,
                      pos_frac, species, order, forces, S(stream));
    });
}

int ofdft_local_pp_stress_pme(ofdft_plan *P, const void *rho, int nspecies, const double *q0, const double *dq,
                              const int *ncoef, const double *const *coef, const double *qmax, const double *vsmall,
                              const double *dqs, int nat, const double *pos_cart, const double *pos_frac,
                              const int *species, int order, double energy, double *stress, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && rho && stress && nat >= 1 && nspecies >= 1, "local_pp_stress_pme: bad arguments");
        for (int a = 0; a < nat; ++a)
            OFDFT_REQUIRE(species[a] >= 0 && species[a] < nspecies,
                          "local_pp_stress_pme: species index out of range");
        pp_stress_pme(*P, (const double *)rho, nspecies, q0, dq, ncoef, coef, qmax, vsmall, dqs, nat, pos_cart,
                      pos_frac, species, order, energy, stress, S(stream));
    });
}

int ofdft_ewald_forces_stress_pme(ofdft_plan *P, int nat, const double *pos_cart, const double *pos_frac,
                                  const double *charges, double precision, int order, double *forces, double *stress,
                                  void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && nat >= 1 && (forces || stress), "ewald_forces_stress_pme: bad arguments");
        if (forces)
            for (int k = 0; k < 3 * nat; ++k) forces[k] = 0.0;
        ewald_fs_pme(*P, nat, pos_cart, pos_frac, charges, precision, order, forces, stress, S(stream));
    });
}

int ofdft_ewald_forces_stress(ofdft_plan *P, int nat, const double *pos_cart, const double *pos_frac,
                              const double *charges, double precision, double *forces, double *stress, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && nat >= 1 && (forces || stress), "ewald_forces_stress: bad arguments");
        if (forces)
            for (int k = 0; k < 3 * nat; ++k) forces[k] = 0.0;
        ewald_fs(*P, nat, pos_cart, pos_frac, charges, precision, forces, stress, S(stream));
    });
}

int ofdft_functional_stress(ofdft_plan *P, const void *rho, const double *energies, double *out, void *stream) {
    return guard([&] {
        OFDFT_REQUIRE(P && rho && energies && out, "functional_stress: bad arguments");
        functional_stress(*P, (const double *)rho, energies, out, S(stream));
    });
}

}  // extern "C"
