// Internal declarations shared by the host API and the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/sem_abi.h"

namespace sem {

void set_error(const char* fmt, ...);

#define SEM_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::sem::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                  \
      return SEM_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

#define SEM_TRY(expr)            \
  do {                           \
    int s_ = (expr);             \
    if (s_ < 0) return s_;       \
  } while (0)

#define SEM_ARG(cond, ...)              \
  do {                                  \
    if (!(cond)) {                      \
      ::sem::set_error(__VA_ARGS__);    \
      return SEM_ERR_ARG;               \
    }                                   \
  } while (0)

constexpr int kMaxP = 16;

// Geometry of the structured local layout shared by every kernel: column-major
// (L = nr*Ex rows, NC = ns*Ey columns); element (ex, ey) node (a, b) sits at
// (ex*nr + a) + L*(ey*ns + b). The fused kernels require nr == ns == P.
struct Layout {
  int P;        // points per direction (nr == ns)
  int Ex, Ey;
  int L, NC;    // rows, columns
  int per_x, per_y;
};

}  // namespace sem

struct sem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* ws = nullptr;          // device workspace, grown on demand
  size_t ws_bytes = 0;
  void* ws2 = nullptr;         // second workspace (mixed-precision solver vectors)
  size_t ws2_bytes = 0;
  void* pinned = nullptr;      // small pinned host buffer for scalar readback
  size_t pinned_bytes = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

struct sem_gsmap {
  int64_t n_local = 0, n_global = 0;
  int32_t* perm = nullptr;     // local indices sorted by global id
  int32_t* seg = nullptr;      // n_global + 1 offsets into perm
  int device = 0;
};

struct sem_mesh {
  int nr = 0, ns = 0, Ex = 0, Ey = 0, per_x = 0, per_y = 0;
  int L = 0, NC = 0;
  int64_t nl = 0, ng = 0;
  int device = 0;
  double* f[SEM_NUM_MESH_FIELDS] = {};
  double *Dr = nullptr, *Ds = nullptr, *zr = nullptr, *wr = nullptr, *zs = nullptr, *ws = nullptr;
  float *Dr32 = nullptr, *Ds32 = nullptr;
  void* small = nullptr;   // the one device block Dr .. Ds32 point into
  std::vector<double> hDr, hDs, hzr, hwr, hzs, hws;
  int32_t* l2g = nullptr;
  sem_gsmap* map = nullptr;
  // interpolation matrices [Jr | Js] from a (nr, ns)-point GLL grid onto this
  // mesh's nodes (advect.jl:66-78), uploaded once per source grid
  struct Interp {
    int nr, ns;
    double* J;
  };
  mutable std::vector<Interp> interp;
  sem::Layout layout() const { return sem::Layout{nr, Ex, Ey, L, NC, per_x, per_y}; }
};

struct sem_operator {
  const sem_mesh* mesh = nullptr;
  int precision = SEM_F64;
  sem_op_desc desc{};
  // factors in operator precision (device); owned[] marks library allocations
  void* G11 = nullptr;
  void* G12 = nullptr;
  void* G22 = nullptr;
  void* B = nullptr;
  void* nu_f = nullptr;
  void* k_f = nullptr;
  bool owned[6] = {false, false, false, false, false, false};
  bool has_mass = false;
  // MFMA operand tables (FP16/BF16 bit patterns) for the contraction matrices
  void* Dmfma = nullptr;
  // single-pass FP16/BF16 operators: G11, G12, G22, B stored as FP16 scaled
  // by 2^e (e chosen per array so the largest magnitude is near 2^14);
  // gscale = 2^-e
  float gscale[4] = {1.f, 1.f, 1.f, 1.f};
  int ax_rows = 0;    // SEM_AX_ROWS at creation (0: the heuristic; -2, -4: rows side by side)
  int ax_fma16 = 0;   // SEM_AX_FMA16=1 at creation
  // iterations of the last fresh solve with this operator: the first launch
  // batch of the next one (repeated solves: time steps, adjoints, FD)
  mutable int last_iters = 0;
  // nr != ns or nr > 16: the element-per-workgroup kernel and the map-based
  // QQ' instead of the fused strip kernels (FP64 only)
  bool generic = false;
};

namespace sem {

int ensure_ws(sem_ctx* ctx, size_t bytes);
int ensure_ws2(sem_ctx* ctx, size_t bytes);
int ensure_pinned(sem_ctx* ctx, size_t bytes);

// ---- kernel launchers (defined in the .hip translation units)
int launch_coords(hipStream_t st, const sem_mesh* m, const double* xe, const double* ye,
                  int deform, double amp);
int launch_geometry(hipStream_t st, sem_mesh* m);
int launch_mult(hipStream_t st, const sem_mesh* m);
// l2g and the QQ' map of a structured mesh from per-direction tables (x:
// global index per local row, first copy, second copy or -1, prefix of copy
// counts; then the same for y)
int launch_numbering(hipStream_t st, const sem_mesh* m, const int32_t* const tabs[8], int nxg,
                     int nyg, int32_t* l2g, int32_t* perm, int32_t* seg);
int launch_mask_fill(hipStream_t st, const Layout& g, int bits, double* M);

// Elementwise epilogue fused into an assembled apply (multigrid smoothing),
// applied to each node copy where its assembled, masked value w = (A u)_i is
// final: kind 1 (a Chebyshev step, u = d): r -= w; d' = c1 d + c2 dinv .* r;
// x += d', with d' stored in y; kind 2: r -= w (y is then scratch); kind 3
// (a residual update that starts the Chebyshev smoother): r -= w;
// d = dinv .* r / theta; x += d, with d a separate buffer (y is scratch).
// kind 0: a plain apply.
struct ApplyEpi {
  int kind = 0;
  void* r = nullptr;
  void* x = nullptr;
  const void* dinv = nullptr;
  double c1 = 0.0, c2 = 0.0;
  void* d = nullptr;
  double theta = 1.0;
};

struct ApplyArgs {
  const void* u;
  void* y;
  const void* G11;
  const void* G12;
  const void* G22;
  const void* B;
  const void* Dr;
  const void* Ds;
  const double* zr;   // GLL nodes and weights (the deflated MFMA forms read them)
  const double* wr;
  const void* nu_f;
  const void* k_f;
  double nu, k;
  int has_mass;
  int bc;
  int assemble;
  int mfma;  // 0: FMA contractions; 1..4: MFMA (FP16, BF16, FP16 x3, BF16 x3)
  // mfma 1, 2: G11, G12, G22, B are FP16 arrays; value = stored * gscale[q]
  float gscale[4];
  // test hooks, read from the environment when the operator is created
  // (SEM_AX_ROWS: element rows swept per workgroup, or -2 / -4 rows side by
  // side; SEM_AX_FMA16=1: FMA contractions
  // for 16x16 blocks instead of the exact FP64/FP32 MFMA forms)
  int rows;
  int fma16;
  ApplyEpi epi;
};
template <typename T>
int launch_apply(hipStream_t st, const Layout& g, const ApplyArgs& a);
// launch arguments of an operator apply y = op(u)
ApplyArgs args_of(const sem_operator* op, const void* u, void* y);
// the apply with a fused epilogue: SEM_ERR_UNSUPPORTED (nothing launched, no
// error recorded) unless operator_epi_ok(op)
bool operator_epi_ok(const sem_operator* op);
int operator_apply_epi(sem_ctx* ctx, const sem_operator* op, const void* u, void* y,
                       const ApplyEpi& e);

// Device-resident PCG scalars (pcg.jl:23-55). t[k & 1] holds t = <r, h> of
// iteration k; done/iters/converged/res are set once, by the first kernel of
// the iteration that finds norm(r, Inf) <= tol or k - 1 == maxiter.
struct PcgState {
  int done;
  int iters;
  int converged;
  int pad;
  double res;
  double t[2];
  // alpha of iteration k at [k & 1]: x += alpha p is applied by the next
  // iteration's apply, and by a finishing pass after the last one
  double alpha[2];
  // relative stop test (tol < 0): |tol| * norm(b, Inf), set by iteration 1
  double tol_eff;
};

struct PcgDev {
  const double* part_prev;  // 2*n3 partials: <r, h> sums then norm(r, Inf) maxima
  int n3;
  double* Q;                // partials of <u, A u> from the apply kernels
  PcgState* state;
  double tol;
  int maxiter;
  int k;                    // 1-based iteration being performed
};

// Launch geometry of the fused apply: nk1 strip workgroups, nk2 seam blocks.
struct AxPlan {
  // CX: elements per strip row; RY: element rows a workgroup processes side
  // by side (1, or 2 for the assembled FMA apply); S: element rows per segment
  int CX, RY, nstrips, S, nsegs, nk1, nk2;
  long nxitems, total;
};
// rows > 0 forces rows swept one after another per workgroup (RY = 1)
AxPlan make_plan(const Layout& g, bool assemble, bool pcg, int rows = 0);

struct PcgLaunch {
  const void* r;  // recursive residual
  const void* d;  // diagonal preconditioner (NULL = identity)
  void* p;        // search direction, updated in place to z + beta p
  void* x;        // the iterate, advanced by the previous alpha p
  PcgDev dev;
};
template <typename T>
int launch_apply_pcg(hipStream_t st, const Layout& g, const ApplyArgs& a, const PcgLaunch& pc);

// Terms of a pointwise combination sum_j c_j (B if w_j) x_j, passed by value.
constexpr int kMaxTerms = 8;
struct CombineTerms {
  const double* x[kMaxTerms];
  double c[kMaxTerms];
  int w[kMaxTerms];
  int n;
};
int launch_combine(hipStream_t st, const sem_mesh* m, const CombineTerms& terms, int bits,
                   double* y);
int launch_dot(hipStream_t st, int64_t n, const double* a, const double* b, const double* w,
               double* part, int nparts, double* out);
int launch_vardiff(hipStream_t st, const sem_mesh* m, const double* nu, double* G11, double* G12,
                   double* G22);
int launch_gprod(hipStream_t st, int64_t n, const double* G11, const double* G12,
                 const double* G22, const double* a, const double* b, double* wr, double* ws);
int launch_pointwise(hipStream_t st, int op, int64_t n, double alpha, const double* a,
                     const double* b, double* y);
int launch_axpby(hipStream_t st, int64_t n, double alpha, const double* x, double beta,
                 const double* y, double* out);


int launch_gsmap(hipStream_t st, const sem_gsmap* map, const double* u, double* y);
// QQ' followed by the Dirichlet mask of bits (u and y may alias)
int launch_gsmap_masked(hipStream_t st, const sem_gsmap* map, const Layout& g, int bits,
                        const double* u, double* y);
int launch_gsmap_masked(hipStream_t st, const sem_gsmap* map, const Layout& g, int bits,
                        const float* u, float* y);
int launch_gsmap_f32(hipStream_t st, const sem_gsmap* map, const float* u, float* y);
int launch_mask(hipStream_t st, int64_t n, const double* M, const double* u, double* y);
int launch_mask_bc(hipStream_t st, const Layout& g, int bits, const double* u, double* y);
int launch_mul(hipStream_t st, int64_t n, const double* a, const double* u, double* y);
int launch_abu(hipStream_t st, int m, int n, const double* u, const double* As, int ma,
               int na, const double* Br, int mb, int nb, int trans, double* tmp, double* y);
int launch_convert_f64_f32(hipStream_t st, int64_t n, const double* x, float* y);
// y = (FP16)(x * scale)
int launch_convert_f64_f16(hipStream_t st, int64_t n, const double* x, double scale, void* y);
int launch_convert_f32_f64(hipStream_t st, int64_t n, const float* x, double* y);
int launch_jacobi_local(hipStream_t st, const sem_operator* op, double* dl);
int launch_inv_masked(hipStream_t st, const Layout& g, int bits, const double* d, double* dinv);

// reductions: results left in device memory at out[0]
int launch_wdot(hipStream_t st, const Layout& g, const double* a, const double* b, double* part,
                int nparts, double* out);
int launch_norm_inf(hipStream_t st, int64_t n, const double* a, double* part, int nparts,
                    double* out);
// acc[0] += scale * sum(a .* b .* w) (w may be NULL), acc in device memory
int launch_dot_acc(hipStream_t st, int64_t n, const double* a, const double* b, const double* w,
                   double scale, double* part, int nparts, double* acc);

int launch_grad(hipStream_t st, const sem_mesh* m, const double* u, double* ux, double* uy);
int launch_grad_t(hipStream_t st, const sem_mesh* m, const double* vx, const double* vy,
                  double* y);
int launch_adjoint_dnu(hipStream_t st, const sem_mesh* m, const double* u, const double* lam,
                       double* g);
int launch_advect(hipStream_t st, const sem_mesh* mV, const sem_mesh* mD, const double* J,
                  const double* T, const double* vx, const double* vy, double* y, int transpose);

// Looks up this GPU's scratch array of the allocation-free PCG (device current)
int init_scratch(int device);

// generic path (k_generic.hip): any nr, ns. Element-local
// nu A u + k B u (+ QQ' and the mask when assemble) with op's factors or the
// mesh's (nu_f, k_f fields optional), and PCG with it
int launch_apply_generic(hipStream_t st, const sem_mesh* m, const double* u, double* y,
                         const double* G11, const double* G12, const double* G22,
                         const double* B, double nu, const double* nu_f, double k,
                         const double* k_f, int has_mass, int assemble, int bits);
int pcg_generic(sem_ctx* ctx, const sem_operator* op, const double* d, const double* b, double* x,
                double tol, int maxiter, sem_pcg_info* info);

// PCG driver (k_pcg.hip)
int pcg_run(sem_ctx* ctx, const sem_operator* op, int precond, const void* d, const void* b,
            void* x, double tol, int maxiter, sem_pcg_info* info);

}  // namespace sem
