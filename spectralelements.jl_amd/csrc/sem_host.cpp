// Host side of libsemhip.so: the C ABI declared in include/sem_abi.h.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <new>

#include "sem_common.h"
#include "sem_l0.h"
#include "_gen/gll_dtables.h"

namespace sem {

namespace {
thread_local std::string g_err;
}

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

int ensure_ws(sem_ctx* ctx, size_t bytes) {
  if (ctx->ws_bytes >= bytes) return SEM_OK;
  if (ctx->ws) {
    SEM_HIP(hipStreamSynchronize(ctx->stream));
    SEM_HIP(hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  SEM_HIP(hipMalloc(&ctx->ws, bytes));
  ctx->ws_bytes = bytes;
  return SEM_OK;
}

int ensure_ws2(sem_ctx* ctx, size_t bytes) {
  if (ctx->ws2_bytes >= bytes) return SEM_OK;
  if (ctx->ws2) {
    SEM_HIP(hipStreamSynchronize(ctx->stream));
    SEM_HIP(hipFree(ctx->ws2));
    ctx->ws2 = nullptr;
    ctx->ws2_bytes = 0;
  }
  SEM_HIP(hipMalloc(&ctx->ws2, bytes));
  ctx->ws2_bytes = bytes;
  return SEM_OK;
}

int ensure_pinned(sem_ctx* ctx, size_t bytes) {
  if (ctx->pinned_bytes >= bytes) return SEM_OK;
  if (ctx->pinned) {
    SEM_HIP(hipStreamSynchronize(ctx->stream));
    SEM_HIP(hipHostFree(ctx->pinned));
    ctx->pinned = nullptr;
  }
  SEM_HIP(hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault));
  ctx->pinned_bytes = bytes;
  return SEM_OK;
}

namespace {

template <typename T>
int upload(hipStream_t st, T** dst, const T* src, size_t n) {
  SEM_HIP(hipMalloc(dst, sizeof(T) * std::max<size_t>(n, 1)));
  SEM_HIP(hipMemcpyAsync(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice, st));
  SEM_HIP(hipStreamSynchronize(st));
  return SEM_OK;
}

int build_gsmap(sem_ctx* ctx, int64_t nl, const int32_t* l2g, int64_t ng, sem_gsmap** out) {
  std::vector<int32_t> seg(ng + 1, 0), perm(nl), fill(ng, 0);
  for (int64_t l = 0; l < nl; ++l) {
    const int32_t g = l2g[l];
    if (g < 0 || g >= ng) {
      set_error("gsmap: global id %d of local %lld outside [0, %lld)", g, (long long)l,
                (long long)ng);
      return SEM_ERR_ARG;
    }
    seg[g + 1]++;
  }
  for (int64_t g = 0; g < ng; ++g) seg[g + 1] += seg[g];
  for (int64_t l = 0; l < nl; ++l) {  // ascending l: stable within a segment
    const int32_t g = l2g[l];
    perm[seg[g] + fill[g]++] = (int32_t)l;
  }
  sem_gsmap* m = new (std::nothrow) sem_gsmap();
  if (!m) {
    set_error("out of host memory");
    return SEM_ERR_ARG;
  }
  m->n_local = nl;
  m->n_global = ng;
  hipGetDevice(&m->device);
  int s = upload(ctx->stream, &m->perm, perm.data(), (size_t)nl);
  if (s == SEM_OK) s = upload(ctx->stream, &m->seg, seg.data(), (size_t)ng + 1);
  if (s != SEM_OK) {
    delete m;
    return s;
  }
  *out = m;
  return SEM_OK;
}

// The fused kernels use compile-time GLL tables; they must equal Mesh.Dr.
template <int P>
bool table_equal(const double* D) {
  for (int i = 0; i < P * P; ++i)
    if (gll_d<P>(i) != D[i]) return false;
  return true;
}

bool gll_table_matches(int P, const double* D) {
  switch (P) {
    case 2: return table_equal<2>(D);
    case 3: return table_equal<3>(D);
    case 4: return table_equal<4>(D);
    case 5: return table_equal<5>(D);
    case 6: return table_equal<6>(D);
    case 7: return table_equal<7>(D);
    case 8: return table_equal<8>(D);
    case 9: return table_equal<9>(D);
    case 10: return table_equal<10>(D);
    case 11: return table_equal<11>(D);
    case 12: return table_equal<12>(D);
    case 13: return table_equal<13>(D);
    case 14: return table_equal<14>(D);
    case 15: return table_equal<15>(D);
    case 16: return table_equal<16>(D);
    default: return false;
  }
}

// The fused strip kernels take nr == ns <= 16 (compile-time D tables, equal
// to Mesh.Dr); other meshes run the generic element kernel.
bool fused_mesh(const sem_mesh* m) {
  return m->nr == m->ns && m->nr >= 2 && m->nr <= kMaxP && gll_table_matches(m->nr, m->hDr.data());
}

// the launch-shape test hooks (environment), read once per operator
void read_ax_hooks(int* rows, int* fma16) {
  const char* r = std::getenv("SEM_AX_ROWS");
  const char* f = std::getenv("SEM_AX_FMA16");
  // rows > 0: rows swept per workgroup; -2 / -4: rows side by side
  *rows = r ? std::atoi(r) : 0;
  if (*rows < 0 && *rows != -2 && *rows != -4) *rows = 0;
  *fma16 = (f && std::atoi(f) == 1) ? 1 : 0;
}

int mfma_mode(int precision) {
  switch (precision) {
    case SEM_F16_MFMA: return 1;
    case SEM_BF16_MFMA: return 2;
    case SEM_F16X3_MFMA: return 3;
    case SEM_BF16X3_MFMA: return 4;
    default: return 0;
  }
}

}  // namespace

ApplyArgs args_of(const sem_operator* op, const void* u, void* y) {
  const sem_mesh* m = op->mesh;
  ApplyArgs a{};
  a.u = u;
  a.y = y;
  a.G11 = op->G11;
  a.G12 = op->G12;
  a.G22 = op->G22;
  a.B = op->B;
  const bool f64 = op->precision == SEM_F64;
  a.Dr = f64 ? (const void*)m->Dr : (const void*)m->Dr32;
  a.Ds = f64 ? (const void*)m->Ds : (const void*)m->Ds32;
  a.zr = m->zr;
  a.wr = m->wr;
  a.nu_f = op->nu_f;
  a.k_f = op->k_f;
  a.nu = op->desc.nu;
  a.k = op->desc.k;
  a.has_mass = op->has_mass ? 1 : 0;
  a.bc = op->desc.assemble ? op->desc.bc_bits : 0;
  a.assemble = op->desc.assemble;
  a.mfma = mfma_mode(op->precision);
  for (int q = 0; q < 4; ++q) a.gscale[q] = op->gscale[q];
  a.rows = op->ax_rows;
  a.fma16 = op->ax_fma16;
  return a;
}

bool operator_epi_ok(const sem_operator* op) {
  if (!op || op->generic || !op->desc.assemble) return false;
  if (op->precision != SEM_F64 && op->precision != SEM_F32) return false;
  const sem_mesh* m = op->mesh;
  if (!(m->nr == 2 || m->nr == 4 || m->nr == 8)) return false;
  return make_plan(m->layout(), true, false, op->ax_rows).RY == 2;
}

int operator_apply_epi(sem_ctx* ctx, const sem_operator* op, const void* u, void* y,
                       const ApplyEpi& e) {
  SEM_ARG(ctx && op && u && y, "NULL argument");
  SEM_ARG(u != y, "fused apply: u and y must not alias");
  if (!operator_epi_ok(op)) return SEM_ERR_UNSUPPORTED;
  ApplyArgs a = args_of(op, u, y);
  a.epi = e;
  if (op->precision == SEM_F64) return launch_apply<double>(ctx->stream, op->mesh->layout(), a);
  return launch_apply<float>(ctx->stream, op->mesh->layout(), a);
}

namespace {

void free_mesh(sem_mesh* m) {
  if (!m) return;
  for (int i = 0; i < SEM_NUM_MESH_FIELDS; ++i)
    if (m->f[i]) hipFree(m->f[i]);
  void* ptrs[] = {m->small, m->l2g};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (m->map) {
    hipFree(m->map->perm);
    hipFree(m->map->seg);
    delete m->map;
  }
  for (auto& it : m->interp) hipFree(it.J);
  delete m;
}

}  // namespace
}  // namespace sem

using namespace sem;

extern "C" {

int sem_abi_version(void) { return SEM_ABI_VERSION; }

const char* sem_last_error(void) { return sem::g_err.c_str(); }

int sem_ctx_create(int device, void* stream, sem_ctx** out) {
  SEM_ARG(out, "sem_ctx_create: out is NULL");
  int ndev = 0;
  SEM_HIP(hipGetDeviceCount(&ndev));
  SEM_ARG(device >= 0 && device < ndev, "sem_ctx_create: device %d not in [0, %d)", device, ndev);
  SEM_HIP(hipSetDevice(device));
  SEM_TRY(init_scratch(device));
  sem_ctx* c = new (std::nothrow) sem_ctx();
  SEM_ARG(c, "out of host memory");
  c->device = device;
  c->stream = static_cast<hipStream_t>(stream);
  SEM_HIP(hipEventCreate(&c->ev0));
  SEM_HIP(hipEventCreate(&c->ev1));
  *out = c;
  return SEM_OK;
}

int sem_ctx_destroy(sem_ctx* ctx) {
  if (!ctx) return SEM_OK;
  hipStreamSynchronize(ctx->stream);
  if (ctx->ws) hipFree(ctx->ws);
  if (ctx->ws2) hipFree(ctx->ws2);
  if (ctx->pinned) hipHostFree(ctx->pinned);
  if (ctx->ev0) hipEventDestroy(ctx->ev0);
  if (ctx->ev1) hipEventDestroy(ctx->ev1);
  delete ctx;
  return SEM_OK;
}

int sem_ctx_set_stream(sem_ctx* ctx, void* stream) {
  SEM_ARG(ctx, "ctx is NULL");
  ctx->stream = static_cast<hipStream_t>(stream);
  return SEM_OK;
}

int sem_ctx_synchronize(sem_ctx* ctx) {
  SEM_ARG(ctx, "ctx is NULL");
  SEM_HIP(hipStreamSynchronize(ctx->stream));
  return SEM_OK;
}

int sem_gausslobatto(int n, double* z, double* w) {
  SEM_ARG(n >= 2 && z && w, "sem_gausslobatto: need n >= 2 and output arrays");
  gll(n, z, w);
  return SEM_OK;
}

int sem_derivmat(int n, const double* x, double* D) {
  SEM_ARG(n >= 1 && x && D, "sem_derivmat: bad arguments");
  derivmat(n, x, D);
  return SEM_OK;
}

int sem_interpmat(int no, const double* xo, int ni, const double* xi, double* J) {
  SEM_ARG(no >= 1 && ni >= 1 && xo && xi && J, "sem_interpmat: bad arguments");
  interpmat(no, xo, ni, xi, J);
  return SEM_OK;
}

int sem_bdfextk(int nt, const double* t, int k, double* a, double* b) {
  SEM_ARG(nt >= 1 && k >= 1 && t && a && b, "sem_bdfextk: bad arguments");
  std::vector<double> u;
  for (int i = 0; i < nt; ++i)
    if (std::find(u.begin(), u.end(), t[i]) == u.end()) u.push_back(t[i]);
  const int m = (int)u.size(), kk = m - 1;
  std::vector<double> D((size_t)m * m), J(std::max(kk, 1));
  derivmat(m, u.data(), D.data());
  if (kk > 0) interpmat(1, u.data(), kk, u.data() + 1, J.data());
  for (int i = 0; i < k; ++i) a[i] = i < kk ? J[i] : 0.0;
  for (int i = 0; i <= k; ++i) b[i] = i < m ? D[(size_t)m * i] : 0.0;
  if (kk == 0) a[0] = 1.0;
  return SEM_OK;
}

int sem_semmesh(int E, int n, double* z, double* w) {
  SEM_ARG(E >= 1 && n >= 2 && z, "sem_semmesh: bad arguments");
  semmesh(E, n, z, w);
  return SEM_OK;
}

int sem_mesh_create(sem_ctx* ctx, int nr, int ns, int Ex, int Ey, int periodic_x, int periodic_y,
                    int deform_kind, double deform_param, sem_mesh** out) {
  SEM_ARG(ctx && out, "sem_mesh_create: ctx/out is NULL");
  SEM_ARG(nr >= 2 && ns >= 2 && nr <= 64 && ns <= 64, "sem_mesh_create: need 2 <= nr, ns <= 64");
  SEM_ARG(Ex >= 1 && Ey >= 1, "sem_mesh_create: need Ex, Ey >= 1");
  SEM_ARG(deform_kind >= SEM_DEFORM_NONE && deform_kind <= SEM_DEFORM_AFFINE01,
          "sem_mesh_create: unknown deform kind %d", deform_kind);
  const int64_t L = (int64_t)nr * Ex, NC = (int64_t)ns * Ey;
  SEM_ARG(L * NC < (int64_t(1) << 31), "sem_mesh_create: %lld local nodes exceed int32 indexing",
          (long long)(L * NC));
  SEM_HIP(hipSetDevice(ctx->device));
  sem_mesh* m = new (std::nothrow) sem_mesh();
  SEM_ARG(m, "out of host memory");
  m->nr = nr;
  m->ns = ns;
  m->Ex = Ex;
  m->Ey = Ey;
  m->per_x = periodic_x ? 1 : 0;
  m->per_y = periodic_y ? 1 : 0;
  m->L = (int)L;
  m->NC = (int)NC;
  m->nl = L * NC;
  m->device = ctx->device;
  const int64_t nxg = (int64_t)Ex * (nr - 1) + (m->per_x ? 0 : 1);
  const int64_t nyg = (int64_t)Ey * (ns - 1) + (m->per_y ? 0 : 1);
  m->ng = nxg * nyg;
  hipStream_t st = ctx->stream;
  int s = SEM_OK;
  auto fail = [&](int code) {
    free_mesh(m);
    return code;
  };
  m->hzr.resize(nr);
  m->hwr.resize(nr);
  m->hzs.resize(ns);
  m->hws.resize(ns);
  gll(nr, m->hzr.data(), m->hwr.data());
  gll(ns, m->hzs.data(), m->hws.data());
  m->hDr.resize((size_t)nr * nr);
  m->hDs.resize((size_t)ns * ns);
  derivmat(nr, m->hzr.data(), m->hDr.data());
  derivmat(ns, m->hzs.data(), m->hDs.data());
  for (int i = 0; i < SEM_NUM_MESH_FIELDS; ++i) {
    if (hipMalloc(&m->f[i], sizeof(double) * m->nl) != hipSuccess) {
      set_error("sem_mesh_create: hipMalloc of %lld doubles failed", (long long)m->nl);
      return fail(SEM_ERR_HIP);
    }
  }
  // the small device arrays (Dr, Ds, the GLL nodes and weights, FP32 Dr, Ds
  // and the 1-D node coordinates semmesh gives) in one block and one copy
  std::vector<double> xe(L), ye(NC);
  semmesh(Ex, nr, xe.data(), nullptr);
  semmesh(Ey, ns, ye.data(), nullptr);
  const std::vector<float> dr32(m->hDr.begin(), m->hDr.end()), ds32(m->hDs.begin(), m->hDs.end());
  struct Piece {
    const void* src;
    size_t bytes;
    size_t off;
  };
  Piece pcs[] = {{m->hDr.data(), sizeof(double) * m->hDr.size(), 0},
                 {m->hDs.data(), sizeof(double) * m->hDs.size(), 0},
                 {m->hzr.data(), sizeof(double) * nr, 0},
                 {m->hwr.data(), sizeof(double) * nr, 0},
                 {m->hzs.data(), sizeof(double) * ns, 0},
                 {m->hws.data(), sizeof(double) * ns, 0},
                 {dr32.data(), sizeof(float) * dr32.size(), 0},
                 {ds32.data(), sizeof(float) * ds32.size(), 0},
                 {xe.data(), sizeof(double) * xe.size(), 0},
                 {ye.data(), sizeof(double) * ye.size(), 0}};
  size_t total = 0;
  for (Piece& pc : pcs) {
    pc.off = total;
    total += (pc.bytes + 255) / 256 * 256;
  }
  std::vector<unsigned char> stage(total, 0);
  for (const Piece& pc : pcs) std::memcpy(stage.data() + pc.off, pc.src, pc.bytes);
  if (hipMalloc(&m->small, total) != hipSuccess) {
    set_error("sem_mesh_create: hipMalloc of %zu bytes failed", total);
    return fail(SEM_ERR_HIP);
  }
  if (hipMemcpyAsync(m->small, stage.data(), total, hipMemcpyHostToDevice, st) != hipSuccess) {
    set_error("sem_mesh_create: copying the node tables failed");
    return fail(SEM_ERR_HIP);
  }
  unsigned char* blk = static_cast<unsigned char*>(m->small);
  m->Dr = reinterpret_cast<double*>(blk + pcs[0].off);
  m->Ds = reinterpret_cast<double*>(blk + pcs[1].off);
  m->zr = reinterpret_cast<double*>(blk + pcs[2].off);
  m->wr = reinterpret_cast<double*>(blk + pcs[3].off);
  m->zs = reinterpret_cast<double*>(blk + pcs[4].off);
  m->ws = reinterpret_cast<double*>(blk + pcs[5].off);
  m->Dr32 = reinterpret_cast<float*>(blk + pcs[6].off);
  m->Ds32 = reinterpret_cast<float*>(blk + pcs[7].off);
  const double* dxe = reinterpret_cast<const double*>(blk + pcs[8].off);
  const double* dye = reinterpret_cast<const double*>(blk + pcs[9].off);
  // coordinates (semmesh + ndgrid + deform) and FP64 geometry on the GPU
  s = launch_coords(st, m, dxe, dye, deform_kind, deform_param);
  if (s == SEM_OK) s = launch_geometry(st, m);
  if (s == SEM_OK) s = launch_mult(st, m);
  hipStreamSynchronize(st);   // also keeps stage alive until the copy is done
  if (s != SEM_OK) return fail(s);
  // local-to-global numbering (semq.jl:6-25 in both directions, mesh.jl:87-89)
  // and the QQ' map, built on the device from per-direction tables: per
  // local row its global index, per global index its one or two copies and
  // the prefix of the copy counts
  std::vector<int32_t> tab;
  int64_t off[9];
  auto axis = [&](int64_t Lr, int n, int per, int64_t ngx) -> bool {
    std::vector<int32_t> gof(Lr), c0(ngx, -1), c1(ngx, -1), pre(ngx);
    for (int64_t i = 0; i < Lr; ++i) {
      int64_t g = (i / n) * (n - 1) + (i % n);
      if (per) g %= ngx;
      gof[i] = (int32_t)g;
      if (c0[g] < 0) c0[g] = (int32_t)i;
      else if (c1[g] < 0) c1[g] = (int32_t)i;
      else return false;
    }
    int32_t acc = 0;
    for (int64_t g = 0; g < ngx; ++g) {
      pre[g] = acc;
      acc += c1[g] >= 0 ? 2 : 1;
    }
    for (auto* v : {&gof, &c0, &c1, &pre}) tab.insert(tab.end(), v->begin(), v->end());
    return true;
  };
  if (!axis(L, nr, m->per_x, nxg) || !axis(NC, ns, m->per_y, nyg)) {
    set_error("sem_mesh_create: a global index with more than two copies along one direction");
    return fail(SEM_ERR_ARG);
  }
  {
    const int64_t sz[8] = {L, nxg, nxg, nxg, NC, nyg, nyg, nyg};
    off[0] = 0;
    for (int q = 0; q < 8; ++q) off[q + 1] = off[q] + sz[q];
  }
  int32_t* dtab = nullptr;
  if ((s = upload(st, &dtab, tab.data(), tab.size())) != SEM_OK) return fail(s);
  const int32_t* tabs[8];
  for (int q = 0; q < 8; ++q) tabs[q] = dtab + off[q];
  sem_gsmap* map = new (std::nothrow) sem_gsmap();
  if (!map) {
    hipFree(dtab);
    set_error("out of host memory");
    return fail(SEM_ERR_ARG);
  }
  map->n_local = m->nl;
  map->n_global = m->ng;
  map->device = ctx->device;
  m->map = map;
  if (hipMalloc(&m->l2g, sizeof(int32_t) * m->nl) != hipSuccess ||
      hipMalloc(&map->perm, sizeof(int32_t) * m->nl) != hipSuccess ||
      hipMalloc(&map->seg, sizeof(int32_t) * (m->ng + 1)) != hipSuccess) {
    hipFree(dtab);
    set_error("sem_mesh_create: hipMalloc of the numbering failed");
    return fail(SEM_ERR_HIP);
  }
  s = launch_numbering(st, m, tabs, (int)nxg, (int)nyg, m->l2g, map->perm, map->seg);
  hipStreamSynchronize(st);
  hipFree(dtab);
  if (s != SEM_OK) return fail(s);
  *out = m;
  return SEM_OK;
}

int sem_mesh_set_coords(sem_ctx* ctx, sem_mesh* m, const double* x, const double* y) {
  SEM_ARG(ctx && m && x && y, "sem_mesh_set_coords: NULL argument");
  hipStream_t st = ctx->stream;
  if (x != m->f[SEM_X])
    SEM_HIP(hipMemcpyAsync(m->f[SEM_X], x, sizeof(double) * m->nl, hipMemcpyDeviceToDevice, st));
  if (y != m->f[SEM_Y])
    SEM_HIP(hipMemcpyAsync(m->f[SEM_Y], y, sizeof(double) * m->nl, hipMemcpyDeviceToDevice, st));
  SEM_TRY(launch_geometry(st, m));
  return SEM_OK;
}

int sem_mesh_destroy(sem_mesh* m) {
  if (m) hipDeviceSynchronize();
  free_mesh(m);
  return SEM_OK;
}

int sem_mesh_sizes(const sem_mesh* m, int* nr, int* ns, int* Ex, int* Ey, int64_t* nl,
                   int64_t* ng) {
  SEM_ARG(m, "mesh is NULL");
  if (nr) *nr = m->nr;
  if (ns) *ns = m->ns;
  if (Ex) *Ex = m->Ex;
  if (Ey) *Ey = m->Ey;
  if (nl) *nl = m->nl;
  if (ng) *ng = m->ng;
  return SEM_OK;
}

int sem_mesh_field(const sem_mesh* m, int which, const double** p) {
  SEM_ARG(m && p, "NULL argument");
  SEM_ARG(which >= 0 && which < SEM_NUM_MESH_FIELDS, "sem_mesh_field: bad field id %d", which);
  *p = m->f[which];
  return SEM_OK;
}

int sem_mesh_dmat(const sem_mesh* m, const double** Dr, const double** Ds) {
  SEM_ARG(m, "mesh is NULL");
  if (Dr) *Dr = m->Dr;
  if (Ds) *Ds = m->Ds;
  return SEM_OK;
}

int sem_mesh_gll(const sem_mesh* m, const double** zr, const double** wr, const double** zs,
                 const double** ws) {
  SEM_ARG(m, "mesh is NULL");
  if (zr) *zr = m->zr;
  if (wr) *wr = m->wr;
  if (zs) *zs = m->zs;
  if (ws) *ws = m->ws;
  return SEM_OK;
}

int sem_mesh_l2g(const sem_mesh* m, const int32_t** p) {
  SEM_ARG(m && p, "NULL argument");
  *p = m->l2g;
  return SEM_OK;
}

int sem_mask_bits(const sem_mesh* m, const char* bc, int* bits) {
  SEM_ARG(m && bc && bits, "NULL argument");
  int b = 0;
  for (int q = 0; q < 4; ++q) {
    SEM_ARG(bc[q] == 'D' || bc[q] == 'N', "sem_mask_bits: bc[%d] = '%c' is not 'D' or 'N'", q,
            bc[q]);
    if (bc[q] == 'D') b |= (1 << q);
  }
  if (m->per_x) b &= ~(SEM_BC_XMIN | SEM_BC_XMAX);
  if (m->per_y) b &= ~(SEM_BC_YMIN | SEM_BC_YMAX);
  *bits = b;
  return SEM_OK;
}

int sem_mask_fill(sem_ctx* ctx, const sem_mesh* m, int bits, double* M) {
  SEM_ARG(ctx && m && M, "NULL argument");
  return launch_mask_fill(ctx->stream, m->layout(), bits, M);
}

int sem_gsmap_create(sem_ctx* ctx, int64_t nl, const int32_t* l2g, int64_t ng, sem_gsmap** out) {
  SEM_ARG(ctx && l2g && out && nl >= 1 && ng >= 1, "sem_gsmap_create: bad arguments");
  SEM_ARG(nl < (int64_t(1) << 31), "sem_gsmap_create: too many local nodes");
  return build_gsmap(ctx, nl, l2g, ng, out);
}

int sem_gsmap_destroy(sem_gsmap* map) {
  if (!map) return SEM_OK;
  hipDeviceSynchronize();
  hipFree(map->perm);
  hipFree(map->seg);
  delete map;
  return SEM_OK;
}

int sem_gsmap_apply(sem_ctx* ctx, const sem_gsmap* map, const double* u, double* y) {
  SEM_ARG(ctx && map && u && y, "NULL argument");
  return launch_gsmap(ctx->stream, map, u, y);
}

int sem_gather_scatter(sem_ctx* ctx, const sem_mesh* m, const double* u, double* y) {
  SEM_ARG(ctx && m && u && y, "NULL argument");
  return launch_gsmap(ctx->stream, m->map, u, y);
}

int sem_gather_scatter_f32(sem_ctx* ctx, const sem_mesh* m, const float* u, float* y) {
  SEM_ARG(ctx && m && u && y, "NULL argument");
  return launch_gsmap_f32(ctx->stream, m->map, u, y);
}

int sem_mask(sem_ctx* ctx, int64_t n, const double* M, const double* u, double* y) {
  SEM_ARG(ctx && u && y && n >= 0, "bad arguments");
  return launch_mask(ctx->stream, n, M, u, y);
}

int sem_mask_bc(sem_ctx* ctx, const sem_mesh* m, int bits, const double* u, double* y) {
  SEM_ARG(ctx && m && u && y, "NULL argument");
  return launch_mask_bc(ctx->stream, m->layout(), bits, u, y);
}

int sem_mass(sem_ctx* ctx, const sem_mesh* m, const double* u, double* y) {
  SEM_ARG(ctx && m && u && y, "NULL argument");
  return launch_mul(ctx->stream, m->nl, m->f[SEM_B], u, y);
}

int sem_abu(sem_ctx* ctx, int m, int n, const double* u, const double* As, int ma, int na,
            const double* Br, int mb, int nb, int trans, double* y) {
  SEM_ARG(ctx && u && y && m >= 1 && n >= 1, "sem_abu: bad arguments");
  const int rnb = trans ? mb : nb, rmb = trans ? nb : mb;
  const int rna = trans ? ma : na;
  if (Br) SEM_ARG(rnb >= 1 && m % rnb == 0, "sem_abu: size(Br,2)=%d does not divide m=%d", rnb, m);
  if (As) SEM_ARG(rna >= 1 && n % rna == 0, "sem_abu: size(As,2)=%d does not divide n=%d", rna, n);
  size_t tmpn = 0;
  if (Br && As) tmpn = (size_t)(m / rnb) * rmb * n;
  double* tmp = nullptr;
  if (tmpn) {
    SEM_TRY(ensure_ws(ctx, sizeof(double) * tmpn));
    tmp = static_cast<double*>(ctx->ws);
  }
  return launch_abu(ctx->stream, m, n, u, As, ma, na, Br, mb, nb, trans, tmp, y);
}

int sem_lapl(sem_ctx* ctx, const sem_mesh* m, const double* u, double* y) {
  return sem_hlmz(ctx, m, u, 1.0, nullptr, 0.0, nullptr, y);
}

int sem_hlmz(sem_ctx* ctx, const sem_mesh* m, const double* u, double nu, const double* nu_f,
             double k, const double* k_f, double* y) {
  SEM_ARG(ctx && m && u && y, "NULL argument");
  if (!fused_mesh(m))
    return launch_apply_generic(ctx->stream, m, u, y, m->f[SEM_G11], m->f[SEM_G12],
                                m->f[SEM_G22], m->f[SEM_B], nu, nu_f, k, k_f,
                                (k != 0.0 || k_f) ? 1 : 0, 0, 0);
  ApplyArgs a{};
  a.u = u;
  a.y = y;
  a.G11 = m->f[SEM_G11];
  a.G12 = m->f[SEM_G12];
  a.G22 = m->f[SEM_G22];
  a.B = m->f[SEM_B];
  a.Dr = m->Dr;
  a.Ds = m->Ds;
  a.nu_f = nu_f;
  a.k_f = k_f;
  a.nu = nu;
  a.k = k;
  a.has_mass = (k != 0.0 || k_f) ? 1 : 0;
  a.bc = 0;
  a.assemble = 0;
  read_ax_hooks(&a.rows, &a.fma16);
  return launch_apply<double>(ctx->stream, m->layout(), a);
}

int sem_operator_create(sem_ctx* ctx, const sem_mesh* m, const sem_op_desc* desc, int precision,
                        sem_operator** out) {
  SEM_ARG(ctx && m && desc && out, "NULL argument");
  SEM_ARG(precision >= SEM_F64 && precision <= SEM_BF16X3_MFMA, "bad precision %d", precision);
  const bool generic = !fused_mesh(m);
  if (generic && precision != SEM_F64) {
    set_error("meshes with nr != ns or nr > %d have FP64 operators only (got nr %d, ns %d)",
              kMaxP, m->nr, m->ns);
    return SEM_ERR_UNSUPPORTED;
  }
  if (precision >= SEM_F16_MFMA) {
    if (m->nr != 16) {
      set_error("MFMA operators need nr == ns == 16 (got %d)", m->nr);
      return SEM_ERR_UNSUPPORTED;
    }
  }
  sem_operator* op = new (std::nothrow) sem_operator();
  SEM_ARG(op, "out of host memory");
  op->mesh = m;
  op->precision = precision;
  op->desc = *desc;
  op->generic = generic;
  op->has_mass = desc->k != 0.0 || desc->k_f;
  read_ax_hooks(&op->ax_rows, &op->ax_fma16);
  const double* src[6] = {desc->G11 ? desc->G11 : m->f[SEM_G11],
                          desc->G12 ? desc->G12 : m->f[SEM_G12],
                          desc->G22 ? desc->G22 : m->f[SEM_G22],
                          desc->B ? desc->B : m->f[SEM_B], desc->nu_f, desc->k_f};
  void** dst[6] = {&op->G11, &op->G12, &op->G22, &op->B, &op->nu_f, &op->k_f};
  const bool half = precision == SEM_F16_MFMA || precision == SEM_BF16_MFMA;
  for (int q = 0; q < 6; ++q) {
    if (!src[q]) continue;
    if (precision == SEM_F64) {
      *dst[q] = const_cast<double*>(src[q]);
    } else if (half && q < 4) {
      // FP16 storage, scaled by 2^e so that the largest magnitude is near 2^14
      int s = ensure_ws(ctx, sizeof(double) * 257);
      double mx = 0.0;
      if (s == SEM_OK) {
        double* part = static_cast<double*>(ctx->ws);
        s = launch_norm_inf(ctx->stream, m->nl, src[q], part, 256, part + 256);
        if (s == SEM_OK && (hipMemcpyAsync(&mx, part + 256, sizeof(double), hipMemcpyDeviceToHost,
                                           ctx->stream) != hipSuccess ||
                            hipStreamSynchronize(ctx->stream) != hipSuccess)) {
          set_error("sem_operator_create: reading the factor range failed");
          s = SEM_ERR_HIP;
        }
      }
      if (s == SEM_OK && !std::isfinite(mx)) {
        set_error("sem_operator_create: non-finite geometric factor");
        s = SEM_ERR_NONFINITE;
      }
      void* h = nullptr;
      if (s == SEM_OK && hipMalloc(&h, sizeof(uint16_t) * m->nl) != hipSuccess) {
        set_error("sem_operator_create: hipMalloc failed");
        s = SEM_ERR_HIP;
      }
      if (s != SEM_OK) {
        sem_operator_destroy(op);
        return s;
      }
      op->owned[q] = true;
      *dst[q] = h;
      const int e = mx > 0.0 ? 14 - std::ilogb(mx) : 0;
      op->gscale[q] = std::ldexp(1.0f, -e);
      s = launch_convert_f64_f16(ctx->stream, m->nl, src[q], std::ldexp(1.0, e), h);
      if (s != SEM_OK) {
        sem_operator_destroy(op);
        return s;
      }
    } else {
      float* f = nullptr;
      if (hipMalloc(&f, sizeof(float) * m->nl) != hipSuccess) {
        set_error("sem_operator_create: hipMalloc failed");
        sem_operator_destroy(op);
        return SEM_ERR_HIP;
      }
      op->owned[q] = true;
      *dst[q] = f;
      int s = launch_convert_f64_f32(ctx->stream, m->nl, src[q], f);
      if (s != SEM_OK) {
        sem_operator_destroy(op);
        return s;
      }
    }
  }
  *out = op;
  return SEM_OK;
}

int sem_operator_destroy(sem_operator* op) {
  if (!op) return SEM_OK;
  void* ps[6] = {op->G11, op->G12, op->G22, op->B, op->nu_f, op->k_f};
  bool any = false;
  for (int q = 0; q < 6; ++q) any = any || op->owned[q];
  if (any || op->Dmfma) hipDeviceSynchronize();
  for (int q = 0; q < 6; ++q)
    if (op->owned[q] && ps[q]) hipFree(ps[q]);
  if (op->Dmfma) hipFree(op->Dmfma);
  delete op;
  return SEM_OK;
}

int sem_operator_precision(const sem_operator* op, int* precision) {
  SEM_ARG(op && precision, "NULL argument");
  *precision = op->precision;
  return SEM_OK;
}

int sem_operator_apply(sem_ctx* ctx, const sem_operator* op, const void* u, void* y) {
  SEM_ARG(ctx && op && u && y, "NULL argument");
  SEM_ARG(u != y, "sem_operator_apply: u and y must not alias");
  if (op->generic)
    return launch_apply_generic(
        ctx->stream, op->mesh, static_cast<const double*>(u), static_cast<double*>(y),
        static_cast<const double*>(op->G11), static_cast<const double*>(op->G12),
        static_cast<const double*>(op->G22), static_cast<const double*>(op->B), op->desc.nu,
        static_cast<const double*>(op->nu_f), op->desc.k, static_cast<const double*>(op->k_f),
        op->has_mass ? 1 : 0, op->desc.assemble, op->desc.assemble ? op->desc.bc_bits : 0);
  const ApplyArgs a = args_of(op, u, y);
  if (op->precision == SEM_F64) return launch_apply<double>(ctx->stream, op->mesh->layout(), a);
  return launch_apply<float>(ctx->stream, op->mesh->layout(), a);
}

int sem_operator_jacobi(sem_ctx* ctx, const sem_operator* op, double* dinv) {
  SEM_ARG(ctx && op && dinv, "NULL argument");
  const sem_mesh* m = op->mesh;
  SEM_TRY(ensure_ws(ctx, sizeof(double) * m->nl));
  double* dl = static_cast<double*>(ctx->ws);
  SEM_TRY(launch_jacobi_local(ctx->stream, op, dl));
  SEM_TRY(launch_gsmap(ctx->stream, m->map, dl, dl));
  SEM_TRY(launch_inv_masked(ctx->stream, m->layout(), op->desc.bc_bits, dl, dinv));
  SEM_HIP(hipStreamSynchronize(ctx->stream));  // dl lives in the shared workspace
  return SEM_OK;
}

int sem_pcg(sem_ctx* ctx, const sem_operator* op, int precond, const void* d, const void* b,
            void* x, double tol, int maxiter, sem_pcg_info* info) {
  SEM_ARG(ctx && op && b && x, "NULL argument");
  SEM_ARG(op->desc.assemble, "sem_pcg needs an assembled operator (desc.assemble = 1)");
  SEM_ARG(precond == SEM_PRECOND_IDENTITY ||
              ((precond == SEM_PRECOND_DIAG || precond == SEM_PRECOND_DIAG_LOCAL) && d),
          "sem_pcg: SEM_PRECOND_DIAG(_LOCAL) needs the diagonal d");
  SEM_ARG(x != b, "sem_pcg: x must not alias b");
  SEM_ARG(precond != SEM_PRECOND_DIAG_LOCAL || op->precision == SEM_F64,
          "sem_pcg: SEM_PRECOND_DIAG_LOCAL needs an SEM_F64 operator");
  SEM_ARG(maxiter >= 0, "sem_pcg: maxiter < 0");
  SEM_ARG(tol == tol, "sem_pcg: tol is NaN");
  return pcg_run(ctx, op, precond, d, b, x, tol, maxiter, info);
}

int sem_wdot(sem_ctx* ctx, const sem_mesh* m, const double* a, const double* b, double* result) {
  SEM_ARG(ctx && m && a && b && result, "NULL argument");
  const int np = 256;
  SEM_TRY(ensure_ws(ctx, sizeof(double) * (np + 1)));
  double* part = static_cast<double*>(ctx->ws);
  if (m->nr == m->ns)
    SEM_TRY(launch_wdot(ctx->stream, m->layout(), a, b, part, np, part + np));
  else
    SEM_TRY(launch_dot(ctx->stream, m->nl, a, b, m->f[SEM_MULT], part, np, part + np));
  SEM_HIP(hipMemcpyAsync(result, part + np, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  SEM_HIP(hipStreamSynchronize(ctx->stream));
  return SEM_OK;
}

int sem_dot(sem_ctx* ctx, int64_t n, const double* a, const double* b, const double* w,
            double* result) {
  SEM_ARG(ctx && a && result && n >= 0, "bad arguments");
  const int np = 256;
  SEM_TRY(ensure_ws(ctx, sizeof(double) * (np + 1)));
  double* part = static_cast<double*>(ctx->ws);
  SEM_TRY(launch_dot(ctx->stream, n, a, b, w, part, np, part + np));
  SEM_HIP(hipMemcpyAsync(result, part + np, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  SEM_HIP(hipStreamSynchronize(ctx->stream));
  return SEM_OK;
}

int sem_dot_acc(sem_ctx* ctx, int64_t n, const double* a, const double* b, const double* w,
                double scale, double* acc) {
  SEM_ARG(ctx && a && acc && n >= 0, "bad arguments");
  const int np = 256;
  SEM_TRY(ensure_ws(ctx, sizeof(double) * np));
  return launch_dot_acc(ctx->stream, n, a, b, w, scale, static_cast<double*>(ctx->ws), np, acc);
}

int sem_combine(sem_ctx* ctx, const sem_mesh* m, int nterms, const double* coef,
                const double* const* x, const int* weighted, int bc_bits, int assemble,
                double* y) {
  SEM_ARG(ctx && m && y, "NULL argument");
  SEM_ARG(nterms >= 0 && nterms <= kMaxTerms, "sem_combine: 0 <= nterms <= 8");
  SEM_ARG(nterms == 0 || (coef && x), "sem_combine: NULL coefficient or term list");
  SEM_ARG(bc_bits < 16, "sem_combine: bc_bits must be < 16 (or negative for no mask)");
  CombineTerms t{};
  t.n = nterms;
  for (int j = 0; j < nterms; ++j) {
    SEM_ARG(x[j], "sem_combine: NULL term");
    t.x[j] = x[j];
    t.c[j] = coef[j];
    t.w[j] = weighted ? weighted[j] : 0;
  }
  SEM_TRY(launch_combine(ctx->stream, m, t, bc_bits, y));
  if (assemble) SEM_TRY(launch_gsmap(ctx->stream, m->map, y, y));
  return SEM_OK;
}

int sem_gprod(sem_ctx* ctx, int64_t n, const double* G11, const double* G12, const double* G22,
              const double* a, const double* b, double* wr, double* ws) {
  SEM_ARG(ctx && G11 && G12 && G22 && a && b && wr && ws && n >= 0, "bad arguments");
  return launch_gprod(ctx->stream, n, G11, G12, G22, a, b, wr, ws);
}

int sem_axpby(sem_ctx* ctx, int64_t n, double alpha, const double* x, double beta,
              const double* y, double* out) {
  SEM_ARG(ctx && x && y && out && n >= 0, "bad arguments");
  return launch_axpby(ctx->stream, n, alpha, x, beta, y, out);
}

int sem_pointwise(sem_ctx* ctx, int op, int64_t n, double alpha, const double* a,
                  const double* b, double* y) {
  SEM_ARG(ctx && a && y && n >= 0, "NULL argument");
  SEM_ARG(op >= SEM_PW_MUL && op <= SEM_PW_SCALE, "sem_pointwise: unknown op %d", op);
  SEM_ARG(op >= SEM_PW_SCALAR_DIV || b, "sem_pointwise: op %d needs b", op);
  return launch_pointwise(ctx->stream, op, n, alpha, a, b, y);
}

int sem_vardiff_factors(sem_ctx* ctx, const sem_mesh* m, const double* nu, double* G11,
                        double* G12, double* G22) {
  SEM_ARG(ctx && m && nu && G11 && G12 && G22, "NULL argument");
  return launch_vardiff(ctx->stream, m, nu, G11, G12, G22);
}

int sem_norm_inf(sem_ctx* ctx, int64_t n, const double* a, double* result) {
  SEM_ARG(ctx && a && result && n >= 0, "bad arguments");
  const int np = 256;
  SEM_TRY(ensure_ws(ctx, sizeof(double) * (np + 1)));
  double* part = static_cast<double*>(ctx->ws);
  SEM_TRY(launch_norm_inf(ctx->stream, n, a, part, np, part + np));
  SEM_HIP(hipMemcpyAsync(result, part + np, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  SEM_HIP(hipStreamSynchronize(ctx->stream));
  return SEM_OK;
}

int sem_grad(sem_ctx* ctx, const sem_mesh* m, const double* u, double* ux, double* uy) {
  SEM_ARG(ctx && m && u && ux && uy, "NULL argument");
  return launch_grad(ctx->stream, m, u, ux, uy);
}

int sem_grad_t(sem_ctx* ctx, const sem_mesh* m, const double* vx, const double* vy, double* y) {
  SEM_ARG(ctx && m && vx && vy && y, "NULL argument");
  return launch_grad_t(ctx->stream, m, vx, vy, y);
}

int sem_adjoint_dnu(sem_ctx* ctx, const sem_mesh* m, const double* u, const double* lam,
                    double* g) {
  SEM_ARG(ctx && m && u && lam && g, "NULL argument");
  return launch_adjoint_dnu(ctx->stream, m, u, lam, g);
}

static int advect_common(sem_ctx* ctx, const sem_mesh* mV, const sem_mesh* mD, const double* T,
                         const double* vx, const double* vy, double* y, int transpose) {
  SEM_ARG(ctx && mV && mD && T && vx && vy && y, "NULL argument");
  SEM_ARG(mV->Ex == mD->Ex && mV->Ey == mD->Ey, "advect: meshes differ in element counts");
  SEM_ARG(mV->nr == mV->ns && mD->nr == mD->ns, "advect: needs nr == ns on both meshes");
  SEM_ARG(mV->nr <= kMaxP && mD->nr <= 24, "advect: supports p <= 16 and dealiased p <= 24");
  const int p = mV->nr, pd = mD->nr;
  double* dJ = nullptr;
  for (auto& it : mD->interp)
    if (it.nr == mV->nr && it.ns == mV->ns) dJ = it.J;
  if (!dJ) {  // first use of this source grid: build and upload once
    std::vector<double> J(2 * (size_t)pd * p);
    interpmat(pd, mD->hzr.data(), p, mV->hzr.data(), J.data());
    interpmat(pd, mD->hzs.data(), p, mV->hzs.data(), J.data() + (size_t)pd * p);
    SEM_HIP(hipMalloc(&dJ, sizeof(double) * J.size()));
    if (hipMemcpy(dJ, J.data(), sizeof(double) * J.size(), hipMemcpyHostToDevice) != hipSuccess) {
      hipFree(dJ);
      set_error("advect: upload of the interpolation matrices failed");
      return SEM_ERR_HIP;
    }
    mD->interp.push_back({mV->nr, mV->ns, dJ});
  }
  SEM_TRY(launch_advect(ctx->stream, mV, mD, dJ, T, vx, vy, y, transpose));
  return SEM_OK;
}

int sem_advect(sem_ctx* ctx, const sem_mesh* mV, const sem_mesh* mD, const double* T,
               const double* vx, const double* vy, double* y) {
  return advect_common(ctx, mV, mD, T, vx, vy, y, 0);
}

int sem_advect_t(sem_ctx* ctx, const sem_mesh* mV, const sem_mesh* mD, const double* w,
                 const double* vx, const double* vy, double* y) {
  return advect_common(ctx, mV, mD, w, vx, vy, y, 1);
}

}  // extern "C"
