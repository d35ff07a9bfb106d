// Internal declarations shared by the libofdft_hip translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ofdft_hip.h"

namespace ofdft {

#define OFDFT_HIP(x)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            throw std::runtime_error(std::string(#x) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

#define OFDFT_REQUIRE(cond, msg)                                  \
    do {                                                          \
        if (!(cond)) throw std::invalid_argument(std::string(msg)); \
    } while (0)

void set_error(const std::string &msg);

// Runtime calls that synchronise with the legacy default stream (blocking
// copies and sets, frees, device synchronisation) fail, and invalidate the
// capture, while another thread captures a stream; the library's setup code
// takes this lock around them and the optimizer's graph capture holds it.
std::recursive_mutex &legacy_lock();
#define OFDFT_LEGACY_LOCK std::lock_guard<std::recursive_mutex> legacy_lk_(legacy_lock())

// Wait for the work queued on a stream by polling it.  The host round trips
// of the solvers are a few microseconds of GPU work each; a blocking wait
// adds an interrupt-driven wake-up to every one of them.
inline void stream_wait(hipStream_t s) {
    hipError_t e;
    while ((e = hipStreamQuery(s)) == hipErrorNotReady) {
    }
    OFDFT_HIP(e);
}

// Gated launches: a kernel given a gate (a device double) returns at once
// when the gate holds 0.  The optimizer's device-driven iteration queues its
// kernels ahead of its control decisions this way (null: always run).
__device__ __forceinline__ bool gate_off(const double *gate) { return gate != nullptr && *gate == 0.0; }

template <typename T> struct cpx_t;
template <> struct cpx_t<float> { using type = float2; };
template <> struct cpx_t<double> { using type = double2; };
template <typename T> using cpx = typename cpx_t<T>::type;

constexpr int kMaxStages = 12;

// Factorisation of one FFT length into Stockham stages.
struct AxisPlan {
    int n;
    int nst;
    int rad[kMaxStages];
    int rmin;
};

// Launch shape of one axis pass (tunable).
struct PassShape {
    int tile;     // pencils (columns pass) or rows (rows pass) per workgroup
    int threads;  // workgroup size
    int lds;      // dynamic LDS bytes
};

// Point-wise prologue applied to real inputs of the R2C rows pass.
enum PrologueOp {
    OP_IDENTITY = 0,
    OP_SQRT_RHOM = 1,    // sqrt(rho > 0 ? rho : 1e-30)                vw.py:34-41
    OP_POW_BETA = 2,     // max(rho, 1e-30)^beta                         wt.py:15-19
    OP_DSQRT = 3,        // drho / (2 sqrt(rhom))      (linear response of sqrt)
    OP_DPOW_BETA = 4,    // beta * rhoc^(beta-1) * drho (linear response of rho^beta)
};

struct Prologue {
    int code[4];
    double beta;
    int b56;          // beta == 5/6: powers by sqrt and cbrt
    const void *aux;  // second input for the linear-response ops (rho)
    const void *phi;  // if set, the input x is replaced by 2 phi x (drho = 2 phi p of a rotation p)
};

struct LmgpState;  // LMGP nonlocal KEDF tables and workspace (lmgp.hip)

// variants of the size-specialised FFT passes the autotuner times: the row
// passes have 3 tiles; the column passes 6, tile var % 3 and radix order var / 3
constexpr int kTileVariants = 3;
constexpr int kColVariants = 6;

// device scalars of a plan (FP64): [0, 256) the solvers' reduction results
// and step scalars, [256, 1024) the state of the device-driven TN iteration
constexpr int kScalSlots = 1024;

// One plan: grid geometry, FFT plans, workspace and functional settings.
struct Plan {
    int nr[3];
    int nzc;
    int nzp;  // padded row pitch of the pipeline spectra (128-byte multiple)
    long long N, Nc, Nfull;
    double lat[9];
    double B[9];  // reciprocal lattice rows b_i (grid.py:405-414)
    double volume, dV;
    int prec;   // OFDFT_FP64 / OFDFT_FP32
    int full;   // full-grid (fftfreq on all axes) semantics
    int max_batch;

    AxisPlan ax[3];
    PassShape rows_r2c, rows_c2r, rows_c2c, cols_y, cols_x, cols_yc, cols_xc;
    void *tw[3];  // device twiddle tables, exp(-2 pi i m / n), plan precision

    // workspace
    void *spec;   // max(4, max_batch) * Nc complex   (also full-grid c2c scratch)
    void *work;   // 4 * N real
    double *red;  // reduction partials
    int red_blocks;
    double *scal;  // device scalars
    void *aux;     // optimizer / propagator fields (allocated at create)
    size_t aux_bytes;
    size_t ws_bytes;
    double *pinned;  // 128 doubles of pinned host memory for scalar read-backs (optimizer, propagator)
    double *pinned_dev;  // its device view (hipHostGetDevicePointer), written by publishing reductions

    // functional settings
    double x = 1.0, y = 1.0, alpha = 5.0 / 6.0, beta = 5.0 / 6.0, rho0 = -1.0, sigma = -1.0;
    unsigned terms = 63;
    const void *vloc = nullptr;
    const void *core = nullptr;  // NLCC core density added to rho in the XC term (N reals) or null
    double eion = 0.0;
    LmgpState *lmgp = nullptr;
    bool rho0_cached_valid = false;
    bool use_fixed = true;  // size-specialised FFT kernels where available
    // tile variants of the size-specialised passes (0 = default tile; set by the autotuner)
    int var_r2c = 0, var_c2r = 0, var_cols[3] = {0, 0, 0};
    // radial table of the TN inner-CG preconditioner (M^-1(|G|) at |G| = i dq,
    // i < tn_pre_n; tn_pre_n = 0: no preconditioner)
    double *tn_pre = nullptr;
    // L-BFGS pair memory of the optimizer: 2 lbfgs_pairs fields (s_i, y_i) of
    // the plan precision, reserved by ofdft_plan_reserve_lbfgs
    void *lbfgs_buf = nullptr;
    int lbfgs_pairs = 0;
    int tn_pre_n = 0, tn_pre_cap = 0;
    double tn_pre_dq = 0.0;
    // staging of the setup routines' host inputs (PP tables, positions): a
    // pinned host buffer and a device buffer, grown on demand and reused
    char *stg_h = nullptr, *stg_d = nullptr;
    size_t stg_cap = 0;
    hipEvent_t stg_evt = nullptr;  // recorded after each staging upload
    // PME B-spline moduli of this grid (x, y, z concatenated) for pme_order
    void *pme_b = nullptr;
    int pme_order = 0;
    // workspace of the force and stress routines (grown on demand)
    void *fs_ws = nullptr;
    size_t fs_cap = 0;
    // gate of the kernels launched for this plan (device double, see
    // gate_off; null outside the optimizer's device-driven iteration)
    const double *gate = nullptr;
    // captured graphs of the device-driven TN iteration (opt.hip)
    struct TnGraphs *tn_graphs = nullptr;
};

// pseudo.hip: release the staging buffers and PME tables of a plan
void plan_setup_free(Plan &P);

int r2c_num_partials(const Plan &P, int batch);
// row pitch of the evaluation pipeline's spectra (128-byte aligned rows when
// every axis has a size-specialised kernel, else nz/2+1)
long long pipe_pitch(const Plan &P);

// fft.hip
void plan_fft_setup(Plan &P);
void plan_fft_free(Plan &P);
// `pitch` is the row pitch (complex elements) of the half spectra; <= 0 means
// the unpadded nz/2+1 layout of the public API
// skip_x leaves out the x-axis pass (forward: the last pass; inverse: the
// first), for pipelines that run it fused with the G-space multiply
template <typename T>
void fft_r2c(Plan &P, const T *in, cpx<T> *out, int batch, long long in_bs, long long out_bs,
             double scale, const Prologue *pro, double *sum_partials, hipStream_t s, long long pitch = 0,
             bool skip_x = false);
template <typename T>
void fft_c2r(Plan &P, const cpx<T> *in, cpx<T> *tmp, T *out, int batch, long long in_bs,
             long long tmp_bs, long long out_bs, double scale, hipStream_t s, long long pitch = 0,
             bool skip_x = false);
bool fixed_size(int n);
template <typename T>
void fft_c2c(Plan &P, const cpx<T> *in, cpx<T> *out, bool inverse, int batch, long long bs, double scale,
             hipStream_t s);
// the x/y passes alone (in place allowed), for fused pipelines
template <typename T>
void fft_cols(Plan &P, int axis, bool full_grid, bool inverse, const cpx<T> *in, cpx<T> *out, int batch,
              long long bs, double scale, hipStream_t s, long long pitch = 0);

// single passes, for timing
template <typename T>
void bench_rows_r2c(Plan &P, const T *in, cpx<T> *out, int batch, const Prologue *pro, hipStream_t s);
template <typename T>
void bench_rows_c2r(Plan &P, const cpx<T> *in, T *out, int batch, hipStream_t s);
// the fused x pass of the evaluation (vW, WT and Hartree spectra) on the
// plan's spectrum buffer; false where the x axis has no fused kernel (eval.hip)
template <typename T> bool bench_xfused(Plan &P, hipStream_t s);
// an evaluation with the energy-density / second-derivative fields (eval.hip)
template <typename T>
void eval_fields(Plan &P, const T *rho, unsigned calc, T *pot, double *energies, T *edens, T *v2, hipStream_t s);
// a stage of the E+V pipeline (0 = fused x pass with the G-space multiply,
// 1 = point-wise post pass), eval.hip
template <typename T> void bench_eval_stage(Plan &P, int stage, const T *rho, T *pot, hipStream_t s);

// autotuner of the tile variants (fft.hip); times[12]: per pass (r2c, c2r,
// x cols, y cols) the mean ms of variants 0..2 (optional)
void plan_autotune(Plan &P, hipStream_t s, float *times);

// reductions (vec.hip)
int reduce_blocks(long long n);
template <typename T>
void dot(Plan &P, const T *a, const T *b, const T *c, long long n, double *out, double scale, hipStream_t s);
// With pub_n > 0 the same launch then copies pub_src[0..pub_n) (which may
// include the outputs) to pub_dst, the device view of pinned host memory: a
// host round trip reads it after the stream completes, without a copy call.
void reduce_final(const double *partials, int nblocks, int nterms, double *out, const double *coef,
                  hipStream_t s, const double *gate = nullptr, const double *pub_src = nullptr,
                  double *pub_dst = nullptr, int pub_n = 0);

// eval.hip
template <typename T>
void eval(Plan &P, const T *rho, unsigned calc, T *pot, double *energies, hipStream_t s);
// analytic Hessian-vector product of the residual (V - mu) phi along p
// (drho = 2 phi p is formed inside the kernels)
template <typename T>
void hvp(Plan &P, const T *rho, const T *phi, const T *p, const T *pot, const T *base_a, const T *base_c,
         const double *mu, T *Ap, int nb, hipStream_t s);
void work_slots(const Plan &P, int *slot);

// lmgp.hip: adds the LMGP nonlocal energy (energies[E_NL], E_SUM) and
// potential (into pot) of rho
template <typename T> void lmgp_eval(Plan &P, const T *rho, T *pot, double *energies, hipStream_t s);
void lmgp_free(Plan &P);

}  // namespace ofdft
