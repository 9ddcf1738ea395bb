// capi.hip -- the extern "C" entry points declared in include/pfgpu.h.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "pf_internal.hpp"

static thread_local char g_err[1024] = "";

void pf_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

bool pf_graphs_allowed() {
  static const bool allowed = [] {
    if (std::getenv("PF_NO_GRAPH")) return false;
    if (std::getenv("ROCPROFILER_LIBRARY_CTOR")) return false;
    const char *pre = std::getenv("LD_PRELOAD");
    if (pre && std::strstr(pre, "rocprofiler")) return false;
    return true;
  }();
  return allowed;
}

namespace {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(const pf_ctx *ctx) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != ctx->device) (void)hipSetDevice(ctx->device);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

template <typename T>
int upload(pf_ctx *ctx, DevBuf<T> &dst, const T *src, size_t count) {
  PF_CHECK(dst.ensure(count));
  if (count) PF_HIP(hipMemcpyAsync(dst.p, src, sizeof(T) * count, hipMemcpyHostToDevice, ctx->stream));
  return PF_OK;
}

template <typename T>
int download(pf_ctx *ctx, T *dst, const T *src, size_t count) {
  if (count) PF_HIP(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

int check_ctx(const pf_ctx *ctx) {
  if (!ctx) {
    pf_set_error("null pf_ctx");
    return PF_INVALID;
  }
  return PF_OK;
}

// host-side validation of an uploaded CRS (1-based, nondecreasing row_ptr, columns in range)
int validate_crs(int64_t n, int64_t nnz, const int64_t *row_ptr, const int32_t *col_idx) {
  if (n < 0 || nnz < 0 || !row_ptr || (nnz > 0 && !col_idx)) {
    pf_set_error("invalid CRS arguments");
    return PF_INVALID;
  }
  if (row_ptr[0] != 1 || row_ptr[n] != nnz + 1) {
    pf_set_error("row_ptr must start at 1 and end at nnz+1");
    return PF_INVALID;
  }
  for (int64_t i = 0; i < n; ++i)
    if (row_ptr[i + 1] < row_ptr[i]) {
      pf_set_error("row_ptr decreases at row %lld", (long long)i + 1);
      return PF_INVALID;
    }
  for (int64_t k = 0; k < nnz; ++k)
    if (col_idx[k] < 1 || (int64_t)col_idx[k] > n) {
      pf_set_error("col_idx(%lld) = %d outside 1..n", (long long)k + 1, col_idx[k]);
      return PF_INVALID;
    }
  return PF_OK;
}

int ensure_work(pf_ctx *ctx, SolverWork **w) {
  if (!*w) PF_CHECK(pf_solver_work_create(w));
  (void)ctx;
  return PF_OK;
}

int set_mesh(pf_ctx *ctx, NodeGraph &g, int64_t nn, int64_t ne, int32_t nne, const double *nodcoord,
             const int32_t *elemnod) {
  if (nn < 0 || ne < 0 || nne < 1 || (ne > 0 && !elemnod)) {
    pf_set_error("invalid mesh arguments");
    return PF_INVALID;
  }
  g.nn = nn;
  g.ne = ne;
  g.nne = nne;
  g.built = g.cols_built = false;
  g.has_coords = nodcoord != nullptr;
  if (nodcoord) PF_CHECK(upload(ctx, g.nodcoord, nodcoord, (size_t)nn * 3));
  PF_CHECK(upload(ctx, g.elemnod, elemnod, (size_t)ne * nne));
  return PF_OK;
}

int64_t nnz_of(const NodeGraph &g, int dof) { return (int64_t)dof * dof * g.nblocks; }

}  // namespace

// ================================================================================== misc
extern "C" const char *pf_last_error(void) { return g_err; }

extern "C" int pf_version(void) { return 10000; }

extern "C" void pf_solve_opts_default(pf_solve_opts *o) {
  std::memset(o, 0, sizeof(*o));
  o->itrmax = 100000;
  o->precision = 64;
  o->er0 = (double)1.0e-10f;  // dble(1.0e-10): REAL(4) literal (FEMSolverClass.f90:70-71)
  o->relative_er = (double)1.0e-10f;
  o->inner_tol = 0.05;
  o->target_relres = 1e-8;
  o->preconditioning = 0;
  o->spmv_kind = 0;
  o->check_every = 32;
  o->use_graph = 1;
  o->debug = 0;
  o->stop_mode = 0;
  o->kernel_timing = 0;
  o->jacobi = 2;
  o->inner_max = 300;
  o->use_x0 = 0;
}

// =============================================================================== context
extern "C" int pf_ctx_create(int device, pf_ctx **out) {
  if (!out) return PF_INVALID;
  *out = nullptr;
  int count = 0;
  PF_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) {
    pf_set_error("device %d not present (%d devices)", device, count);
    return PF_INVALID;
  }
  pf_ctx *c = new pf_ctx();
  c->device = device;
  DeviceGuard guard(c);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) {
    pf_set_error("context setup failed: %s", hipGetErrorString(e));
    delete c;
    return PF_ERR_HIP;
  }
  // the device address of the allocation-free solver's scratch array, once per device
  const int rc = pf_noalloc_setup(device);
  if (rc != PF_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return PF_OK;
}

extern "C" int pf_ctx_destroy(pf_ctx *ctx) {
  if (!ctx) return PF_OK;
  {
    DeviceGuard guard(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    pf_solver_work_destroy(ctx->work);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
  return PF_OK;
}

extern "C" int pf_ctx_synchronize(pf_ctx *ctx) {
  PF_CHECK(check_ctx(ctx));
  DeviceGuard guard(ctx);
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

extern "C" void *pf_ctx_stream(pf_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

// ====================================================================== stateless calls
extern "C" int pf_crs_graph(pf_ctx *ctx, int64_t nn, int64_t ne, int32_t nne, const int32_t *elemnod, int32_t dof,
                            int64_t *nnz, int64_t *row_ptr, int32_t *col_idx) {
  PF_CHECK(check_ctx(ctx));
  if (dof < 1 || !nnz) {
    pf_set_error("pf_crs_graph: dof >= 1 and nnz required");
    return PF_INVALID;
  }
  DeviceGuard guard(ctx);
  NodeGraph g;
  PF_CHECK(set_mesh(ctx, g, nn, ne, nne, nullptr, elemnod));
  PF_CHECK(pf_graph_node_cols(ctx, g));
  *nnz = nnz_of(g, dof);
  if (!row_ptr && !col_idx) return PF_OK;
  const int64_t n = (int64_t)dof * nn;
  DevBuf<int64_t> rp;
  DevBuf<int32_t> ci;
  PF_CHECK(rp.ensure((size_t)n + 1));
  PF_CHECK(ci.ensure((size_t)*nnz + 1));
  PF_CHECK(pf_graph_expand(ctx, g, dof, rp.p, ci.p));
  if (row_ptr) PF_CHECK(download(ctx, row_ptr, rp.p, (size_t)n + 1));
  if (col_idx) PF_CHECK(download(ctx, col_idx, ci.p, (size_t)*nnz));
  return PF_OK;
}

extern "C" int pf_hex8_stiffness(pf_ctx *ctx, int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                                 const double *E, const double *nu, int32_t jmode, int64_t first, int64_t count,
                                 double *ke) {
  PF_CHECK(check_ctx(ctx));
  if (first < 1 || count < 0 || first - 1 + count > ne || !E || !nu || !ke || !nodcoord) {
    pf_set_error("pf_hex8_stiffness: element range %lld..%lld outside 1..%lld", (long long)first,
                 (long long)(first + count - 1), (long long)ne);
    return PF_INVALID;
  }
  DeviceGuard guard(ctx);
  NodeGraph g;
  PF_CHECK(set_mesh(ctx, g, nn, ne, 8, nodcoord, elemnod));
  PF_CHECK(pf_graph_check_mesh(ctx, g));
  DevBuf<double> dE, dnu, dke;
  PF_CHECK(upload(ctx, dE, E, (size_t)ne));
  PF_CHECK(upload(ctx, dnu, nu, (size_t)ne));
  PF_CHECK(dke.ensure((size_t)count * 576));
  PF_CHECK(pf_hex8_ke_batch(ctx, g, dE.p, dnu.p, jmode, first - 1, count, dke.p));
  return download(ctx, ke, dke.p, (size_t)count * 576);
}

extern "C" int pf_hex8_mass_vector(pf_ctx *ctx, int64_t nn, int64_t ne, const double *nodcoord,
                                   const int32_t *elemnod, double density, const double *accel, int64_t first,
                                   int64_t count, double *mv) {
  PF_CHECK(check_ctx(ctx));
  if (first < 1 || count < 0 || first - 1 + count > ne || !accel || !mv || !nodcoord) {
    pf_set_error("pf_hex8_mass_vector: element range %lld..%lld outside 1..%lld or null argument",
                 (long long)first, (long long)(first + count - 1), (long long)ne);
    return PF_INVALID;
  }
  DeviceGuard guard(ctx);
  NodeGraph g;
  PF_CHECK(set_mesh(ctx, g, nn, ne, 8, nodcoord, elemnod));
  PF_CHECK(pf_graph_check_mesh(ctx, g));
  DevBuf<double> dmv;
  PF_CHECK(dmv.ensure((size_t)count * 24 + 1));
  PF_CHECK(pf_hex8_mass_batch(ctx, g, density, accel, first - 1, count, dmv.p));
  return download(ctx, mv, dmv.p, (size_t)count * 24);
}

extern "C" int pf_hex8_mass_matrix(pf_ctx *ctx, int64_t nn, int64_t ne, const double *nodcoord,
                                   const int32_t *elemnod, const double *density, int64_t ndensity, int64_t first,
                                   int64_t count, double *me) {
  PF_CHECK(check_ctx(ctx));
  if (first < 1 || count < 0 || first - 1 + count > ne || !me || !nodcoord || !density ||
      !(ndensity == 1 || ndensity == ne)) {
    pf_set_error("pf_hex8_mass_matrix: element range %lld..%lld outside 1..%lld, a null argument, or a density "
                 "count other than 1 or ne", (long long)first, (long long)(first + count - 1), (long long)ne);
    return PF_INVALID;
  }
  DeviceGuard guard(ctx);
  NodeGraph g;
  PF_CHECK(set_mesh(ctx, g, nn, ne, 8, nodcoord, elemnod));
  PF_CHECK(pf_graph_check_mesh(ctx, g));
  DevBuf<double> drho, dme;
  PF_CHECK(upload(ctx, drho, density, (size_t)ndensity));
  PF_CHECK(dme.ensure((size_t)count * 576 + 1));
  PF_CHECK(pf_hex8_mass_mat(ctx, g, drho.p, ndensity == ne ? 1 : 0, first - 1, count, dme.p));
  return download(ctx, me, dme.p, (size_t)count * 576);
}

static int stress_params(int64_t ne, int64_t nE, int64_t nnu, const double *E, const double *nu) {
  if (!E || !nu || !(nE == 1 || nE == ne) || !(nnu == 1 || nnu == ne)) {
    pf_set_error("stress: size(E) and size(v) must be 1 or the number of elements");
    return PF_INVALID;
  }
  return PF_OK;
}

extern "C" int pf_element_stress(pf_ctx *ctx, int64_t nn, int64_t ne, const double *nodcoord, const int32_t *elemnod,
                                 const double *disp, const double *E, int64_t nE, const double *nu, int64_t nnu,
                                 int32_t jmode, double *sigma) {
  PF_CHECK(check_ctx(ctx));
  PF_CHECK(stress_params(ne, nE, nnu, E, nu));
  if (!disp || !sigma || !nodcoord) {
    pf_set_error("pf_element_stress: null argument");
    return PF_INVALID;
  }
  DeviceGuard guard(ctx);
  NodeGraph g;
  PF_CHECK(set_mesh(ctx, g, nn, ne, 8, nodcoord, elemnod));
  PF_CHECK(pf_graph_check_mesh(ctx, g));
  DevBuf<double> du, dE, dnu, ds;
  PF_CHECK(upload(ctx, du, disp, (size_t)(3 * nn)));
  PF_CHECK(upload(ctx, dE, E, (size_t)nE));
  PF_CHECK(upload(ctx, dnu, nu, (size_t)nnu));
  PF_CHECK(ds.ensure((size_t)ne * 9 + 1));
  PF_CHECK(pf_cell_stress(ctx, g, du.p, dE.p, nE == ne && ne > 1 ? 1 : 0, dnu.p, nnu == ne && ne > 1 ? 1 : 0,
                          jmode, ds.p));
  return download(ctx, sigma, ds.p, (size_t)ne * 9);
}

extern "C" int pf_hex8_stiffness_crs(pf_ctx *ctx, int64_t nn, int64_t ne, const double *nodcoord,
                                     const int32_t *elemnod, const double *E, const double *nu, int32_t jmode,
                                     const int64_t *row_ptr, const int32_t *col_idx, double *val) {
  PF_CHECK(check_ctx(ctx));
  if (!E || !nu || !row_ptr || !col_idx || !val || !nodcoord) {
    pf_set_error("pf_hex8_stiffness_crs: null argument");
    return PF_INVALID;
  }
  DeviceGuard guard(ctx);
  NodeGraph g;
  PF_CHECK(set_mesh(ctx, g, nn, ne, 8, nodcoord, elemnod));
  PF_CHECK(pf_graph_node_cols(ctx, g));
  const int64_t n = 3 * nn, nnz = nnz_of(g, 3);
  DevBuf<int64_t> rp;
  DevBuf<int32_t> ci;
  DevBuf<double> dv, dE, dnu;
  PF_CHECK(rp.ensure((size_t)n + 1));
  PF_CHECK(ci.ensure((size_t)nnz + 1));
  PF_CHECK(pf_graph_expand(ctx, g, 3, rp.p, ci.p));
  // the given pattern must be this mesh's ZeroMatrix(DOF=3)
  std::vector<int64_t> hrp((size_t)n + 1);
  std::vector<int32_t> hci((size_t)nnz);
  PF_CHECK(download(ctx, hrp.data(), rp.p, (size_t)n + 1));
  PF_CHECK(download(ctx, hci.data(), ci.p, (size_t)nnz));
  if (std::memcmp(hrp.data(), row_ptr, sizeof(int64_t) * (n + 1)) != 0 ||
      std::memcmp(hci.data(), col_idx, sizeof(int32_t) * nnz) != 0) {
    pf_set_error("pf_hex8_stiffness_crs: row_ptr/col_idx are not ZeroMatrix(DOF=3) of this mesh");
    return PF_INVALID;
  }
  PF_CHECK(upload(ctx, dE, E, (size_t)ne));
  PF_CHECK(upload(ctx, dnu, nu, (size_t)ne));
  PF_CHECK(dv.ensure((size_t)nnz + 1));
  PF_CHECK(pf_assemble_hex8(ctx, g, dE.p, dnu.p, jmode, rp.p, dv.p));
  return download(ctx, val, dv.p, (size_t)nnz);
}

extern "C" int pf_spmv_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                           const double *x, double *y) {
  PF_CHECK(check_ctx(ctx));
  if (n < 0 || !row_ptr) return PF_INVALID;
  const int64_t nnz = row_ptr[n] - 1;
  PF_CHECK(validate_crs(n, nnz, row_ptr, col_idx));
  DeviceGuard guard(ctx);
  DevBuf<int64_t> rp;
  DevBuf<int32_t> ci;
  DevBuf<double> dv, dx, dy;
  PF_CHECK(upload(ctx, rp, row_ptr, (size_t)n + 1));
  PF_CHECK(upload(ctx, ci, col_idx, (size_t)nnz));
  PF_CHECK(upload(ctx, dv, val, (size_t)nnz));
  PF_CHECK(upload(ctx, dx, x, (size_t)n));
  PF_CHECK(dy.ensure((size_t)n + 1));
  PF_CHECK(pf_spmv_crs_launch<double>(ctx, n, rp.p, ci.p, dv.p, dx.p, dy.p, n ? (double)nnz / n : 1.0));
  return download(ctx, y, dy.p, (size_t)n);
}

static void fill_stats(pf_stats *st, const SolveResult &r) {
  if (!st) return;
  std::memset(st, 0, sizeof(*st));
  st->spmv_p_ms = r.t_ms[0];
  st->spmv_e_ms = r.t_ms[1];
  st->update_ms = r.t_ms[2];
  st->pupdate_ms = r.t_ms[3];
  st->timed_iterations = r.timed;
  st->status = r.status;
  st->iterations = r.iterations;
  st->outer_iterations = r.outer;
  st->spmv_kind_used = r.kind;
  st->rr = r.rr;
  st->rr1 = r.rr1;
  st->true_relres = r.true_relres;
  st->solve_ms = r.solve_ms;
  st->setup_ms = r.setup_ms;
}

extern "C" int pf_bicgstab_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                               const double *val, const double *rhs, double *x, const pf_solve_opts *opts,
                               pf_stats *stats) {
  PF_CHECK(check_ctx(ctx));
  if (n < 0 || !row_ptr || !rhs || !x) return PF_INVALID;
  const int64_t nnz = row_ptr[n] - 1;
  PF_CHECK(validate_crs(n, nnz, row_ptr, col_idx));
  pf_solve_opts o;
  if (opts) o = *opts; else pf_solve_opts_default(&o);
  DeviceGuard guard(ctx);
  PF_CHECK(ensure_work(ctx, &ctx->work));
  DevBuf<int64_t> rp;
  DevBuf<int32_t> ci;
  DevBuf<double> dv, db, dx;
  PF_CHECK(upload(ctx, rp, row_ptr, (size_t)n + 1));
  PF_CHECK(upload(ctx, ci, col_idx, (size_t)nnz));
  PF_CHECK(upload(ctx, dv, val, (size_t)nnz));
  PF_CHECK(upload(ctx, db, rhs, (size_t)n));
  PF_CHECK(upload(ctx, dx, x, (size_t)n));
  if (o.preconditioning < 0 || o.preconditioning > 3) {
    pf_set_error("preconditioning must be 0 (none), 1 (PointJacobi), 2 (ILU) or 3 (PointJacobi, no solve)");
    return PF_INVALID;
  }
  if (o.preconditioning == 1 || o.preconditioning == 3) PF_CHECK(pf_jacobi_rows(ctx, n, rp.p, ci.p, dv.p, db.p));
  SolveResult res;
  if (o.preconditioning == 3) {
    fill_stats(stats, res);
    return PF_OK;
  }
  PF_CHECK(pf_bicgstab_device(ctx, ctx->work, n, rp.p, ci.p, dv.p, nnz, db.p, dx.p, o, true, res));
  fill_stats(stats, res);
  PF_CHECK(download(ctx, x, dx.p, (size_t)n));
  // graphs captured on these temporary buffers must not be replayed later
  pf_solver_work_destroy(ctx->work);
  ctx->work = nullptr;
  return res.status;
}

extern "C" int pf_jacobi_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, double *val,
                             double *rhs) {
  PF_CHECK(check_ctx(ctx));
  if (n < 0 || !row_ptr || !val) return PF_INVALID;
  const int64_t nnz = row_ptr[n] - 1;
  PF_CHECK(validate_crs(n, nnz, row_ptr, col_idx));
  DeviceGuard guard(ctx);
  DevBuf<int64_t> rp;
  DevBuf<int32_t> ci;
  DevBuf<double> dv, db;
  PF_CHECK(upload(ctx, rp, row_ptr, (size_t)n + 1));
  PF_CHECK(upload(ctx, ci, col_idx, (size_t)nnz));
  PF_CHECK(upload(ctx, dv, val, (size_t)nnz));
  if (rhs) PF_CHECK(upload(ctx, db, rhs, (size_t)n));
  PF_CHECK(pf_jacobi_rows(ctx, n, rp.p, ci.p, dv.p, rhs ? db.p : nullptr));
  PF_CHECK(download(ctx, val, dv.p, (size_t)nnz));
  if (rhs) PF_CHECK(download(ctx, rhs, db.p, (size_t)n));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

namespace {
// node-block structure of a CRS on the device (ILU groups of three rows when present)
int ilu_group(pf_ctx *ctx, int64_t n, const int64_t *rp, const int32_t *ci, int *G) {
  DevBuf<int32_t> flag;
  PF_CHECK(flag.ensure(1));
  int ok = 0;
  PF_CHECK(pf_block_structure_check(ctx, n, rp, ci, n / 3, flag.p, &ok));
  *G = ok ? 3 : 1;
  return PF_OK;
}
}  // namespace

extern "C" int pf_ilu0_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *val,
                           double *fval) {
  PF_CHECK(check_ctx(ctx));
  if (n < 0 || !row_ptr || !val || !fval) return PF_INVALID;
  const int64_t nnz = row_ptr[n] - 1;
  PF_CHECK(validate_crs(n, nnz, row_ptr, col_idx));
  DeviceGuard guard(ctx);
  DevBuf<int64_t> rp;
  DevBuf<int32_t> ci;
  DevBuf<double> dv;
  PF_CHECK(upload(ctx, rp, row_ptr, (size_t)n + 1));
  PF_CHECK(upload(ctx, ci, col_idx, (size_t)nnz));
  PF_CHECK(upload(ctx, dv, val, (size_t)nnz));
  int G = 1;
  PF_CHECK(ilu_group(ctx, n, rp.p, ci.p, &G));
  IluFactor F;
  PF_CHECK(pf_ilu_factor(ctx, n, rp.p, ci.p, dv.p, nnz, G, F));
  PF_CHECK(download(ctx, fval, F.f.p, (size_t)nnz));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

extern "C" int pf_ilu_apply_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                                const double *fval, const double *v, double *y) {
  PF_CHECK(check_ctx(ctx));
  if (n < 0 || !row_ptr || !fval || !v || !y) return PF_INVALID;
  const int64_t nnz = row_ptr[n] - 1;
  PF_CHECK(validate_crs(n, nnz, row_ptr, col_idx));
  DeviceGuard guard(ctx);
  DevBuf<int64_t> rp;
  DevBuf<int32_t> ci;
  DevBuf<double> df, dvv, dy;
  PF_CHECK(upload(ctx, rp, row_ptr, (size_t)n + 1));
  PF_CHECK(upload(ctx, ci, col_idx, (size_t)nnz));
  PF_CHECK(upload(ctx, df, fval, (size_t)nnz));
  PF_CHECK(upload(ctx, dvv, v, (size_t)n));
  PF_CHECK(dy.ensure((size_t)n + 1));
  int G = 1;
  PF_CHECK(ilu_group(ctx, n, rp.p, ci.p, &G));
  IluFactor F;
  PF_CHECK(pf_ilu_factor(ctx, n, rp.p, ci.p, df.p, nnz, G, F, false));
  IluSweepArgs a;
  a.v = dvv.p;
  a.out = dy.p;
  PF_CHECK(pf_ilu_sweep_enqueue(ctx, F, 0, a));
  PF_CHECK(download(ctx, y, dy.p, (size_t)n));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

extern "C" int pf_dev_spmv_crs(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                               const double *val, const double *x, double *y) {
  PF_CHECK(check_ctx(ctx));
  if (n <= 0) return n == 0 ? PF_OK : PF_INVALID;
  DeviceGuard guard(ctx);
  int64_t end = 0;
  PF_CHECK(pf_read_i64(ctx, row_ptr + n, &end));
  PF_CHECK(pf_spmv_crs_launch<double>(ctx, n, row_ptr, col_idx, val, x, y, (double)(end - 1) / n));
  return PF_OK;
}

extern "C" int pf_dev_spmv_crs_f32(pf_ctx *ctx, int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                                   const float *val, const float *x, float *y) {
  PF_CHECK(check_ctx(ctx));
  if (n <= 0) return n == 0 ? PF_OK : PF_INVALID;
  DeviceGuard guard(ctx);
  int64_t end = 0;
  PF_CHECK(pf_read_i64(ctx, row_ptr + n, &end));
  PF_CHECK(pf_spmv_crs_launch<float>(ctx, n, row_ptr, col_idx, val, x, y, (double)(end - 1) / n));
  return PF_OK;
}

namespace {
// new_vector += A old_vector on host arrays, as crs_spmv_real64/real32 (SparseClass-c.c:38-66)
template <typename V>
int crs_spmv_acc_host(pf_ctx *ctx, const V *val, const int64_t *row_ptr, const int32_t *col_idx,
                      const V *old_vector, V *new_vector, int64_t n, int64_t col_size) {
  PF_CHECK(check_ctx(ctx));
  if (n < 0 || col_size < 0 || !row_ptr) {
    pf_set_error("crs_spmv: invalid sizes or null row_ptr");
    return PF_INVALID;
  }
  if (n == 0) return PF_OK;
  if (!old_vector || !new_vector) {
    pf_set_error("crs_spmv: null vector");
    return PF_INVALID;
  }
  const int64_t nnz = row_ptr[n] - 1;
  PF_CHECK(validate_crs(n, nnz, row_ptr, col_idx));
  if (nnz > 0 && !val) {
    pf_set_error("crs_spmv: null val");
    return PF_INVALID;
  }
  if (nnz > col_size) {
    pf_set_error("crs_spmv: row_ptr(n+1)-1 = %lld entries but col_size = %lld", (long long)nnz,
                 (long long)col_size);
    return PF_INVALID;
  }
  DeviceGuard guard(ctx);
  DevBuf<int64_t> rp;
  DevBuf<int32_t> ci;
  DevBuf<V> dv, dx, dy;
  PF_CHECK(upload(ctx, rp, row_ptr, (size_t)n + 1));
  PF_CHECK(upload(ctx, ci, col_idx, (size_t)nnz));
  PF_CHECK(upload(ctx, dv, val, (size_t)nnz));
  PF_CHECK(upload(ctx, dx, old_vector, (size_t)n));
  PF_CHECK(upload(ctx, dy, (const V *)new_vector, (size_t)n));
  PF_CHECK(pf_spmv_crs_launch<V>(ctx, n, rp.p, ci.p, dv.p, dx.p, dy.p, (double)nnz / (double)n, true));
  return download(ctx, new_vector, dy.p, (size_t)n);
}

// the context of the reference-named exports: device 0, created on first use
pf_ctx *reference_ctx() {
  static pf_ctx *ctx = [] {
    pf_ctx *c = nullptr;
    if (pf_ctx_create(0, &c) != PF_OK) c = nullptr;
    return c;
  }();
  return ctx;
}

template <typename V>
void crs_spmv_reference_abi(const char *name, const V *val, const long *row_ptr, const int *col_idx,
                            const V *old_vector, V *new_vector, size_t n, size_t col_size) {
  static_assert(sizeof(long) == sizeof(int64_t) && sizeof(int) == sizeof(int32_t), "LP64 ABI expected");
  pf_ctx *ctx = reference_ctx();
  const int rc = ctx ? crs_spmv_acc_host<V>(ctx, val, (const int64_t *)row_ptr, (const int32_t *)col_idx,
                                            old_vector, new_vector, (int64_t)n, (int64_t)col_size)
                     : PF_ERR_HIP;
  if (rc != PF_OK) {
    // the reference's signature returns nothing: a failed product must not pass silently
    std::fprintf(stderr, "%s: %s\n", name, ctx ? pf_last_error() : "no HIP device");
    std::abort();
  }
}
}  // namespace

extern "C" int pf_crs_spmv_real64(pf_ctx *ctx, const double *val, const int64_t *row_ptr, const int32_t *col_idx,
                                  const double *old_vector, double *new_vector, int64_t n, int64_t col_size) {
  return crs_spmv_acc_host<double>(ctx, val, row_ptr, col_idx, old_vector, new_vector, n, col_size);
}

extern "C" int pf_crs_spmv_real32(pf_ctx *ctx, const float *val, const int64_t *row_ptr, const int32_t *col_idx,
                                  const float *old_vector, float *new_vector, int64_t n, int64_t col_size) {
  return crs_spmv_acc_host<float>(ctx, val, row_ptr, col_idx, old_vector, new_vector, n, col_size);
}

extern "C" void crs_spmv_real64(double val[], long row_ptr[], int col_idx[], double old_vector[],
                                double new_vector[], size_t n, size_t col_size) {
  crs_spmv_reference_abi<double>("crs_spmv_real64", val, row_ptr, col_idx, old_vector, new_vector, n, col_size);
}

extern "C" void crs_spmv_real32(float val[], long row_ptr[], int col_idx[], float old_vector[],
                                float new_vector[], size_t n, size_t col_size) {
  crs_spmv_reference_abi<float>("crs_spmv_real32", val, row_ptr, col_idx, old_vector, new_vector, n, col_size);
}

// ============================================================================ pf_system
struct pf_system {
  pf_ctx *ctx = nullptr;
  NodeGraph g;
  bool has_mesh = false;
  int dof = 0;
  int64_t n = 0, nnz = 0;
  bool graph_crs = false;  // CRS is ZeroMatrix(dof) of the mesh
  bool multi_crs = false;  // CRS is the multi-domain pattern (node blocks, columns ascending)
  DevBuf<int64_t> row_ptr;
  DevBuf<int32_t> col_idx;
  DevBuf<double> val, rhs, x, ubar;
  DevBuf<uint8_t> fixed;
  bool has_fix = false;
  bool matrix_changed = true;
  SolverWork *work = nullptr;
  DevBuf<double> sx, sy;  // pf_system_spmv operands
  DevBuf<double> E, nu;   // resident materials
  bool has_mat = false;
  uint64_t mat_version = 1;  // bumped with every change of the CRS values or structure
  uint64_t rhs_version = 1, fix_version = 1;
  uint64_t bc_mat = 0, bc_rhs = 0, bc_fix = 0;  // versions the elimination was last applied to
  BatchWork *batch = nullptr;
  MultiWork *multi = nullptr;
  DirectWork *direct = nullptr;
  uint64_t pattern_version = 1;  // bumped with every change of the CRS structure
  int32_t direct_mfma = 1;       // direct solve: 1 MFMA tile products, 0 the VALU form
  int32_t ke_kernel = 0;         // element matrices: 0 reference order, 1/2 isotropic split VALU/MFMA
  int32_t jmode = 0;          // J mode of the last assembly (reused by the sensitivities)
  bool ubar_nonzero = false;  // some fixed DOF has a nonzero prescribed value
  // keepThisMatrixAs: named copies of the CRS values (same pattern) and the fix_eig list
  struct Kept {
    std::string name;
    int64_t nnz = 0;
    DevBuf<double> val;
  };
  std::vector<std::unique_ptr<Kept>> kept;
  std::vector<int32_t> fix_eig;
  ~pf_system() {
    pf_solver_work_destroy(work);
    pf_batch_work_destroy(batch);
    pf_multi_work_destroy(multi);
    pf_direct_work_destroy(direct);
  }
};

static inline void mark_changed(pf_system *s) {
  s->matrix_changed = true;
  ++s->mat_version;
}

// Dirichlet elimination (FEMSolverClass.f90:1707-1756) of the current fixes on the current
// CRS and RHS, skipped when none of them changed since it last ran (it is idempotent on the
// values it already produced, so repeated solves keep their cached layouts)
static int apply_bc(pf_system *s) {
  if (!s->has_fix) return PF_OK;
  if (s->bc_mat == s->mat_version && s->bc_rhs == s->rhs_version && s->bc_fix == s->fix_version) return PF_OK;
  const bool matrix_part = s->bc_mat != s->mat_version || s->bc_fix != s->fix_version;
  PF_CHECK(pf_bc_apply(s->ctx, s->n, s->row_ptr.p, s->col_idx.p, s->val.p, s->rhs.p, s->fixed.p, s->ubar.p,
                       s->n > 0 ? (double)s->nnz / (double)s->n : 1.0));
  if (matrix_part) mark_changed(s);
  s->bc_mat = s->mat_version;
  s->bc_rhs = s->rhs_version;
  s->bc_fix = s->fix_version;
  return PF_OK;
}

namespace {
int sys_check(pf_system *s, bool need_crs) {
  if (!s || !s->ctx) {
    pf_set_error("null pf_system");
    return PF_INVALID;
  }
  if (need_crs && s->n == 0 && s->nnz == 0 && !s->row_ptr.p) {
    pf_set_error("no CRS: call pf_system_zero_matrix or pf_system_set_crs first");
    return PF_INVALID;
  }
  return PF_OK;
}

int sys_alloc_vectors(pf_system *s, bool zero) {
  pf_ctx *ctx = s->ctx;
  PF_CHECK(s->rhs.ensure((size_t)s->n + 1));
  PF_CHECK(s->x.ensure((size_t)s->n + 1));
  if (zero) {
    PF_HIP(hipMemsetAsync(s->rhs.p, 0, sizeof(double) * (s->n + 1), ctx->stream));
    PF_HIP(hipMemsetAsync(s->x.p, 0, sizeof(double) * (s->n + 1), ctx->stream));
  }
  s->has_fix = false;
  s->ubar_nonzero = false;
  ++s->rhs_version;
  ++s->fix_version;
  return PF_OK;
}
}  // namespace

extern "C" int pf_system_create(pf_ctx *ctx, pf_system **out) {
  PF_CHECK(check_ctx(ctx));
  if (!out) return PF_INVALID;
  pf_system *s = new pf_system();
  s->ctx = ctx;
  *out = s;
  return PF_OK;
}

extern "C" int pf_system_destroy(pf_system *s) {
  if (!s) return PF_OK;
  DeviceGuard guard(s->ctx);
  (void)hipStreamSynchronize(s->ctx->stream);
  delete s;
  return PF_OK;
}

extern "C" int pf_system_set_mesh(pf_system *s, int64_t nn, int64_t ne, int32_t nne, const double *nodcoord,
                                  const int32_t *elemnod) {
  PF_CHECK(sys_check(s, false));
  DeviceGuard guard(s->ctx);
  PF_CHECK(set_mesh(s->ctx, s->g, nn, ne, nne, nodcoord, elemnod));
  PF_CHECK(pf_graph_check_mesh(s->ctx, s->g));
  s->has_mesh = true;
  s->has_mat = false;
  s->graph_crs = false;
  s->multi_crs = false;
  ++s->pattern_version;
  return PF_OK;
}

extern "C" int pf_system_zero_matrix(pf_system *s, int32_t dof, int64_t *nnz) {
  PF_CHECK(sys_check(s, false));
  if (!s->has_mesh || dof < 1) {
    pf_set_error("pf_system_zero_matrix: mesh required and dof >= 1");
    return PF_INVALID;
  }
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  PF_CHECK(pf_graph_node_cols(ctx, s->g));
  s->dof = dof;
  s->n = (int64_t)dof * s->g.nn;
  s->nnz = nnz_of(s->g, dof);
  if (s->nnz >= (int64_t)1 << 40) {
    pf_set_error("CRS too large");
    return PF_INVALID;
  }
  PF_CHECK(s->row_ptr.ensure((size_t)s->n + 1));
  PF_CHECK(s->col_idx.ensure((size_t)s->nnz + 1));
  PF_CHECK(s->val.ensure((size_t)s->nnz + 1));
  PF_CHECK(pf_graph_expand(ctx, s->g, dof, s->row_ptr.p, s->col_idx.p));
  PF_HIP(hipMemsetAsync(s->val.p, 0, sizeof(double) * (s->nnz + 1), ctx->stream));
  PF_CHECK(sys_alloc_vectors(s, true));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  s->graph_crs = true;
  s->multi_crs = false;
  ++s->pattern_version;
  mark_changed(s);
  if (nnz) *nnz = s->nnz;
  return PF_OK;
}

extern "C" int pf_system_set_crs(pf_system *s, int64_t n, int64_t nnz, const int64_t *row_ptr,
                                 const int32_t *col_idx, const double *val) {
  PF_CHECK(sys_check(s, false));
  PF_CHECK(validate_crs(n, nnz, row_ptr, col_idx));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  const bool resize = (n != s->n);
  s->n = n;
  s->nnz = nnz;
  s->dof = 0;
  PF_CHECK(upload(ctx, s->row_ptr, row_ptr, (size_t)n + 1));
  PF_CHECK(upload(ctx, s->col_idx, col_idx, (size_t)nnz));
  if (val) {
    PF_CHECK(upload(ctx, s->val, val, (size_t)nnz));
  } else {
    PF_CHECK(s->val.ensure((size_t)nnz + 1));
    PF_HIP(hipMemsetAsync(s->val.p, 0, sizeof(double) * (nnz + 1), ctx->stream));
  }
  if (resize || !s->rhs.p) PF_CHECK(sys_alloc_vectors(s, true));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  s->graph_crs = false;
  s->multi_crs = false;
  ++s->pattern_version;
  mark_changed(s);
  return PF_OK;
}

extern "C" int pf_system_sizes(pf_system *s, int64_t *n, int64_t *nnz, int64_t *nn, int64_t *ne) {
  PF_CHECK(sys_check(s, false));
  if (n) *n = s->n;
  if (nnz) *nnz = s->nnz;
  if (nn) *nn = s->g.nn;
  if (ne) *ne = s->g.ne;
  return PF_OK;
}

extern "C" int pf_system_get_crs(pf_system *s, int64_t *row_ptr, int32_t *col_idx, double *val) {
  PF_CHECK(sys_check(s, true));
  DeviceGuard guard(s->ctx);
  if (row_ptr) PF_CHECK(download(s->ctx, row_ptr, s->row_ptr.p, (size_t)s->n + 1));
  if (col_idx) PF_CHECK(download(s->ctx, col_idx, s->col_idx.p, (size_t)s->nnz));
  if (val) PF_CHECK(download(s->ctx, val, s->val.p, (size_t)s->nnz));
  return PF_OK;
}

// stiffness on the multi-domain pattern: values cleared, every entry summed over the
// incident elements of the concatenated mesh in ascending id (the order of setMatrix calls
// looping over domains and their elements)
static int multi_stiffness(pf_system *s, int32_t jmode) {
  pf_ctx *ctx = s->ctx;
  if (!s->g.built) PF_CHECK(pf_graph_node_elems(ctx, s->g));
  DevBuf<int32_t> err;
  PF_CHECK(err.ensure(1));
  PF_HIP(hipMemsetAsync(err.p, 0, sizeof(int32_t), ctx->stream));
  PF_HIP(hipMemsetAsync(s->val.p, 0, sizeof(double) * (s->nnz + 1), ctx->stream));
  PF_CHECK(pf_assemble_hex8_search(ctx, s->g, s->E.p, s->nu.p, jmode, s->row_ptr.p, s->col_idx.p, true, s->val.p,
                                   err.p));
  int32_t herr = 0;
  PF_CHECK(download(ctx, &herr, err.p, 1));
  if (herr) {
    pf_set_error("stiffness: no CRS slot for an element entry (the multi-domain CRS is empty or stale)");
    return PF_INVALID;
  }
  return PF_OK;
}

extern "C" int pf_system_stiffness(pf_system *s, const double *E, const double *nu, int32_t jmode) {
  PF_CHECK(sys_check(s, false));
  if (!s->has_mesh || s->g.nne != 8 || (!E != !nu) || (!E && !s->has_mat)) {
    pf_set_error("pf_system_stiffness: hex8 mesh and E, nu (or resident ones) required");
    return PF_INVALID;
  }
  if (!(s->graph_crs && s->dof == 3) && !(s->multi_crs && s->dof == 3))
    PF_CHECK(pf_system_zero_matrix(s, 3, nullptr));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  if (E) {
    PF_CHECK(upload(ctx, s->E, E, (size_t)s->g.ne));
    PF_CHECK(upload(ctx, s->nu, nu, (size_t)s->g.ne));
    s->has_mat = true;
  }
  if (s->multi_crs) PF_CHECK(multi_stiffness(s, jmode));
  else PF_CHECK(pf_assemble_hex8(ctx, s->g, s->E.p, s->nu.p, jmode, s->row_ptr.p, s->val.p, s->ke_kernel));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  s->jmode = jmode;
  mark_changed(s);
  return PF_OK;
}

extern "C" int pf_system_stiffness_dev(pf_system *s, const double *E_dev, const double *nu_dev, int32_t jmode) {
  PF_CHECK(sys_check(s, false));
  if (!s->has_mesh || s->g.nne != 8 || !E_dev || !nu_dev) {
    pf_set_error("pf_system_stiffness_dev: hex8 mesh and device arrays E, nu required");
    return PF_INVALID;
  }
  if (!(s->graph_crs && s->dof == 3) && !(s->multi_crs && s->dof == 3))
    PF_CHECK(pf_system_zero_matrix(s, 3, nullptr));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  PF_CHECK(s->E.ensure((size_t)s->g.ne));
  PF_CHECK(s->nu.ensure((size_t)s->g.ne));
  PF_HIP(hipMemcpyAsync(s->E.p, E_dev, sizeof(double) * s->g.ne, hipMemcpyDeviceToDevice, ctx->stream));
  PF_HIP(hipMemcpyAsync(s->nu.p, nu_dev, sizeof(double) * s->g.ne, hipMemcpyDeviceToDevice, ctx->stream));
  s->has_mat = true;
  if (s->multi_crs) PF_CHECK(multi_stiffness(s, jmode));
  else PF_CHECK(pf_assemble_hex8(ctx, s->g, s->E.p, s->nu.p, jmode, s->row_ptr.p, s->val.p, s->ke_kernel));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  s->jmode = jmode;
  mark_changed(s);
  return PF_OK;
}

static int check_elem_ids(const pf_system *s, int64_t count, const int32_t *ids) {
  for (int64_t q = 0; q < count; ++q)
    if (ids[q] < 1 || (int64_t)ids[q] > s->g.ne) {
      pf_set_error("element id %d outside 1..%lld", ids[q], (long long)s->g.ne);
      return PF_INVALID;
    }
  return PF_OK;
}

extern "C" int pf_system_add_element_matrices(pf_system *s, int64_t count, const int32_t *elem_ids,
                                              const double *ke) {
  PF_CHECK(sys_check(s, true));
  if (!s->has_mesh || s->dof < 1 || count < 0 || (count > 0 && (!elem_ids || !ke))) {
    pf_set_error("pf_system_add_element_matrices: needs a mesh and a CRS from pf_system_zero_matrix");
    return PF_INVALID;
  }
  if (count == 0) return PF_OK;
  PF_CHECK(check_elem_ids(s, count, elem_ids));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  const int64_t nk = (int64_t)s->g.nne * s->dof;
  DevBuf<int32_t> dids, err;
  DevBuf<double> dke;
  PF_CHECK(upload(ctx, dids, elem_ids, (size_t)count));
  PF_CHECK(upload(ctx, dke, ke, (size_t)(count * nk * nk)));
  PF_CHECK(err.ensure(1));
  PF_HIP(hipMemsetAsync(err.p, 0, sizeof(int32_t), ctx->stream));
  if (!s->g.built) PF_CHECK(pf_graph_node_elems(ctx, s->g));
  PF_CHECK(pf_add_element_mats(ctx, s->g, s->dof, count, dids.p, dke.p, s->row_ptr.p, s->col_idx.p, s->val.p,
                               err.p));
  int32_t herr = 0;
  PF_CHECK(download(ctx, &herr, err.p, 1));
  mark_changed(s);
  if (herr) {
    pf_set_error("setMatrix: fill-in request, no CRS slot for an element entry");
    return PF_INVALID;
  }
  return PF_OK;
}

extern "C" int pf_system_set_rhs(pf_system *s, const double *rhs) {
  PF_CHECK(sys_check(s, true));
  DeviceGuard guard(s->ctx);
  PF_CHECK(upload(s->ctx, s->rhs, rhs, (size_t)s->n));
  PF_HIP(hipStreamSynchronize(s->ctx->stream));
  ++s->rhs_version;
  return PF_OK;
}

extern "C" int pf_system_add_element_vectors(pf_system *s, int64_t count, const int32_t *elem_ids,
                                             const double *vec) {
  PF_CHECK(sys_check(s, true));
  if (!s->has_mesh || s->dof < 1 || count < 0 || (count > 0 && (!elem_ids || !vec))) {
    pf_set_error("pf_system_add_element_vectors: needs a mesh and a CRS from pf_system_zero_matrix");
    return PF_INVALID;
  }
  if (count == 0) return PF_OK;
  PF_CHECK(check_elem_ids(s, count, elem_ids));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  const int64_t nk = (int64_t)s->g.nne * s->dof;
  DevBuf<int32_t> dids;
  DevBuf<double> dvec;
  PF_CHECK(upload(ctx, dids, elem_ids, (size_t)count));
  PF_CHECK(upload(ctx, dvec, vec, (size_t)(count * nk)));
  if (!s->g.built) PF_CHECK(pf_graph_node_elems(ctx, s->g));
  ++s->rhs_version;
  return pf_add_element_vecs(ctx, s->g, s->dof, count, dids.p, dvec.p, s->rhs.p);
}

extern "C" int pf_system_get_rhs(pf_system *s, double *rhs) {
  PF_CHECK(sys_check(s, true));
  DeviceGuard guard(s->ctx);
  return download(s->ctx, rhs, s->rhs.p, (size_t)s->n);
}

extern "C" int pf_system_add_traction(pf_system *s, const double *box, double pressure) {
  PF_CHECK(sys_check(s, true));
  if (!s->has_mesh || s->g.nne != 8 || s->dof != 3 || !box) {
    pf_set_error("pf_system_add_traction: hex8 mesh with a DOF=3 CRS required");
    return PF_INVALID;
  }
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  if (!s->g.built) PF_CHECK(pf_graph_node_elems(ctx, s->g));
  PF_CHECK(pf_traction_box(ctx, s->g, box, pressure, s->rhs.p));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  ++s->rhs_version;
  return PF_OK;
}

extern "C" int pf_system_fix(pf_system *s, int64_t nids, const int32_t *ids, const double *values,
                             double fixvalue) {
  PF_CHECK(sys_check(s, true));
  if (nids < 0 || (nids > 0 && !ids)) return PF_INVALID;
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  if (!s->has_fix) {
    PF_CHECK(s->fixed.ensure((size_t)s->n + 1));
    PF_CHECK(s->ubar.ensure((size_t)s->n + 1));
    PF_HIP(hipMemsetAsync(s->fixed.p, 0, s->n + 1, ctx->stream));
    PF_HIP(hipMemsetAsync(s->ubar.p, 0, sizeof(double) * (s->n + 1), ctx->stream));
    s->has_fix = true;
  }
  ++s->fix_version;
  if (nids == 0) return PF_OK;
  if (values) {
    for (int64_t q = 0; q < nids; ++q)
      if (values[q] != 0.0 && ids[q] >= 1 && ids[q] <= s->n) s->ubar_nonzero = true;
  } else if (fixvalue != 0.0) {
    s->ubar_nonzero = true;
  }
  DevBuf<int32_t> dids;
  DevBuf<double> dvals;
  PF_CHECK(upload(ctx, dids, ids, (size_t)nids));
  if (values) PF_CHECK(upload(ctx, dvals, values, (size_t)nids));
  return pf_bc_fix(ctx, s->n, nids, dids.p, values ? dvals.p : nullptr, fixvalue, s->fixed.p, s->ubar.p);
}

extern "C" int pf_system_solve(pf_system *s, const pf_solve_opts *opts, const double *x0, double *x_out,
                               pf_stats *stats) {
  PF_CHECK(sys_check(s, true));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  pf_solve_opts o;
  if (opts) o = *opts; else pf_solve_opts_default(&o);
  PF_CHECK(ensure_work(ctx, &s->work));
  if (o.preconditioning < 0 || o.preconditioning > 3) {
    pf_set_error("preconditioning must be 0 (none), 1 (PointJacobi), 2 (ILU) or 3 (PointJacobi, no solve)");
    return PF_INVALID;
  }
  PF_CHECK(apply_bc(s));
  if (o.preconditioning == 1 || o.preconditioning == 3) {
    // JacobiPreconditionerCRS on CRS_val / CRS_RHS after the elimination
    // (FEMSolverClass.f90:1769-1775); the elimination stays valid (the fixed rows hold
    // a unit diagonal), re-running it is a no-op
    PF_CHECK(pf_jacobi_rows(ctx, s->n, s->row_ptr.p, s->col_idx.p, s->val.p, s->rhs.p));
    mark_changed(s);
    ++s->rhs_version;
  }
  if (x0) {
    PF_CHECK(upload(ctx, s->x, x0, (size_t)s->n));
  } else {
    PF_HIP(hipMemsetAsync(s->x.p, 0, sizeof(double) * (s->n + 1), ctx->stream));
  }
  SolveResult res;
  if (o.preconditioning == 3) {
    // the reference's PointJacobi branch calls no solver: x = x0 (or zeros), then the fixed values
    res.status = PF_OK;
  } else {
    PF_CHECK(pf_bicgstab_device(ctx, s->work, s->n, s->row_ptr.p, s->col_idx.p, s->val.p, s->nnz, s->rhs.p,
                                s->x.p, o, s->matrix_changed, res));
    s->matrix_changed = false;
  }
  if (s->has_fix) PF_CHECK(pf_bc_set_fixed_x(ctx, s->n, s->fixed.p, s->ubar.p, s->x.p));
  fill_stats(stats, res);
  if (x_out) PF_CHECK(download(ctx, x_out, s->x.p, (size_t)s->n));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return res.status;
}

extern "C" int pf_system_ilu_factor(pf_system *s, double *fval, int32_t *levels) {
  PF_CHECK(sys_check(s, true));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  PF_CHECK(ensure_work(ctx, &s->work));
  PF_CHECK(apply_bc(s));
  IluFactor *F = nullptr;
  PF_CHECK(pf_work_ilu(ctx, s->work, s->n, s->row_ptr.p, s->col_idx.p, s->val.p, s->nnz, s->matrix_changed, &F));
  s->matrix_changed = false;
  if (levels) *levels = (int32_t)F->lev.size() - 1;
  if (fval) PF_CHECK(download(ctx, fval, F->f.p, (size_t)s->nnz));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

extern "C" int pf_system_ilu_apply(pf_system *s, const double *v, double *y) {
  PF_CHECK(sys_check(s, true));
  if (!v || !y) return PF_INVALID;
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  PF_CHECK(pf_system_ilu_factor(s, nullptr, nullptr));
  IluFactor *F = nullptr;
  PF_CHECK(pf_work_ilu(ctx, s->work, s->n, s->row_ptr.p, s->col_idx.p, s->val.p, s->nnz, false, &F));
  PF_CHECK(s->sx.ensure((size_t)s->n + 1));
  PF_CHECK(s->sy.ensure((size_t)s->n + 1));
  PF_CHECK(upload(ctx, s->sx, v, (size_t)s->n));
  IluSweepArgs a;
  a.v = s->sx.p;
  a.out = s->sy.p;
  PF_CHECK(pf_ilu_sweep_enqueue(ctx, *F, 0, a));
  PF_CHECK(download(ctx, y, s->sy.p, (size_t)s->n));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

extern "C" int pf_system_ilu_reset(pf_system *s) {
  PF_CHECK(sys_check(s, false));
  pf_work_ilu_reset(s->work);
  return PF_OK;
}

extern "C" int pf_system_mass_matrix(pf_system *s, const double *density, double density_all) {
  PF_CHECK(sys_check(s, false));
  if (!s->has_mesh || s->g.nne != 8) {
    pf_set_error("pf_system_mass_matrix: hex8 mesh required");
    return PF_INVALID;
  }
  if (!(s->graph_crs && s->dof == 3)) PF_CHECK(pf_system_zero_matrix(s, 3, nullptr));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  DevBuf<double> drho;
  if (density) PF_CHECK(upload(ctx, drho, density, (size_t)s->g.ne));
  else PF_CHECK(upload(ctx, drho, &density_all, 1));
  PF_CHECK(pf_assemble_mass(ctx, s->g, drho.p, density ? 1 : 0, s->row_ptr.p, s->val.p));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  mark_changed(s);
  return PF_OK;
}

extern "C" int pf_system_zeros(pf_system *s) {
  PF_CHECK(sys_check(s, true));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  PF_HIP(hipMemsetAsync(s->val.p, 0, sizeof(double) * (s->nnz + 1), ctx->stream));
  if (s->rhs.p) PF_HIP(hipMemsetAsync(s->rhs.p, 0, sizeof(double) * (s->n + 1), ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  mark_changed(s);
  ++s->rhs_version;
  return PF_OK;
}

extern "C" int pf_system_keep_matrix(pf_system *s, const char *name) {
  PF_CHECK(sys_check(s, true));
  if (!name) return PF_INVALID;
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  pf_system::Kept *k = nullptr;
  for (auto &e : s->kept)
    if (e->name == name) k = e.get();
  if (!k) {
    s->kept.emplace_back(new pf_system::Kept());
    k = s->kept.back().get();
    k->name = name;
  }
  PF_CHECK(k->val.ensure((size_t)s->nnz + 1));
  PF_HIP(hipMemcpyAsync(k->val.p, s->val.p, sizeof(double) * s->nnz, hipMemcpyDeviceToDevice, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  k->nnz = s->nnz;
  return PF_OK;
}

extern "C" int pf_system_fix_eig(pf_system *s, int64_t nids, const int32_t *ids) {
  PF_CHECK(sys_check(s, false));
  if (nids < 0 || (nids > 0 && !ids)) return PF_INVALID;
  s->fix_eig.insert(s->fix_eig.end(), ids, ids + nids);
  std::sort(s->fix_eig.begin(), s->fix_eig.end());
  s->fix_eig.erase(std::unique(s->fix_eig.begin(), s->fix_eig.end()), s->fix_eig.end());
  return PF_OK;
}

extern "C" int pf_system_eig_count(pf_system *s, int32_t *count) {
  PF_CHECK(sys_check(s, true));
  if (!count) return PF_INVALID;
  int64_t nfix = 0;
  for (int32_t id : s->fix_eig)
    if (id >= 1 && id <= s->n) ++nfix;
  *count = (int32_t)(s->n - nfix);
  return PF_OK;
}

extern "C" int pf_system_eig(pf_system *s, int32_t nev, double tol, int32_t maxit, double *eigen_values,
                             double *eigen_vectors, int32_t *count, pf_stats *stats) {
  PF_CHECK(sys_check(s, true));
  pf_system::Kept *A = nullptr, *B = nullptr;
  for (auto &e : s->kept) {
    if (e->name == "A") A = e.get();
    if (e->name == "B") B = e.get();
  }
  if (!A || !B || A->nnz != s->nnz || B->nnz != s->nnz) {
    pf_set_error("eig: keep the stiffness as \"A\" and the mass as \"B\" on the current CRS pattern first");
    return PF_INVALID;
  }
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  int32_t cnt = 0;
  EigResult r;
  PF_CHECK(pf_eig_device(ctx, s->n, s->row_ptr.p, s->col_idx.p, A->val.p, B->val.p, s->nnz, s->fix_eig, nev,
                         tol > 0.0 ? tol : 1e-8, maxit > 0 ? maxit : 10000, eigen_values, eigen_vectors, &cnt, r));
  if (count) *count = cnt;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->status = r.status;
    stats->iterations = r.iterations;
    stats->true_relres = r.max_res;
    stats->solve_ms = r.ms;
  }
  return r.status;
}

extern "C" int pf_system_get_x(pf_system *s, double *x) {
  PF_CHECK(sys_check(s, true));
  DeviceGuard guard(s->ctx);
  return download(s->ctx, x, s->x.p, (size_t)s->n);
}

extern "C" int pf_system_view_get(pf_system *s, pf_system_view *v) {
  PF_CHECK(sys_check(s, false));
  if (!v) return PF_INVALID;
  v->n = s->n;
  v->nnz = s->nnz;
  v->nn = s->g.nn;
  v->ne = s->g.ne;
  v->row_ptr = s->row_ptr.p;
  v->col_idx = s->col_idx.p;
  v->val = s->val.p;
  v->rhs = s->rhs.p;
  v->x = s->x.p;
  v->nodcoord = s->g.nodcoord.p;
  v->elemnod = s->g.elemnod.p;
  return PF_OK;
}

extern "C" int pf_system_spmv(pf_system *s, int32_t kind, const double *x, double *y, int32_t repeat, double *ms) {
  PF_CHECK(sys_check(s, true));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  PF_CHECK(ensure_work(ctx, &s->work));
  PF_CHECK(s->sy.ensure((size_t)s->n + 1));
  if (!s->sx.p || x) {
    PF_CHECK(s->sx.ensure((size_t)s->n + 1));
    if (x) PF_CHECK(upload(ctx, s->sx, x, (size_t)s->n));
    else PF_HIP(hipMemsetAsync(s->sx.p, 0, sizeof(double) * (s->n + 1), ctx->stream));
  }
  PF_CHECK(pf_work_spmv(ctx, s->work, s->n, s->row_ptr.p, s->col_idx.p, s->val.p, s->nnz, s->matrix_changed, kind,
                        s->sx.p, s->sy.p, repeat, ms));
  s->matrix_changed = false;
  if (y) PF_CHECK(download(ctx, y, s->sy.p, (size_t)s->n));
  return PF_OK;
}

// ============================================================================ batch
extern "C" int pf_system_add_tractions(pf_system *s, int32_t nmesh, const int64_t *node_off, const double *boxes,
                                       const double *pressures) {
  PF_CHECK(sys_check(s, true));
  if (!s->has_mesh || s->g.nne != 8 || s->dof != 3 || nmesh < 0 || (nmesh > 0 && (!node_off || !boxes || !pressures))) {
    pf_set_error("pf_system_add_tractions: hex8 mesh with a DOF=3 CRS, node offsets, boxes and pressures required");
    return PF_INVALID;
  }
  if (nmesh == 0) return PF_OK;
  for (int32_t m = 0; m < nmesh; ++m)
    if (node_off[m] < 0 || node_off[m + 1] < node_off[m] || node_off[m + 1] > s->g.nn) {
      pf_set_error("pf_system_add_tractions: node offsets must be non-decreasing within 0..%lld", (long long)s->g.nn);
      return PF_INVALID;
    }
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  if (!s->g.built) PF_CHECK(pf_graph_node_elems(ctx, s->g));
  DevBuf<int64_t> doff;
  DevBuf<double> dbox, dp;
  PF_CHECK(upload(ctx, doff, node_off, (size_t)nmesh + 1));
  PF_CHECK(upload(ctx, dbox, boxes, (size_t)nmesh * 6));
  PF_CHECK(upload(ctx, dp, pressures, (size_t)nmesh));
  PF_CHECK(pf_traction_boxes(ctx, s->g, nmesh, doff.p, dbox.p, dp.p, s->rhs.p));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  ++s->rhs_version;
  return PF_OK;
}

extern "C" int pf_system_batch_set(pf_system *s, int32_t nmesh, const int64_t *node_off, const double *cost) {
  PF_CHECK(sys_check(s, true));
  DeviceGuard guard(s->ctx);
  return pf_batch_configure(s->ctx, &s->batch, s->n, s->row_ptr.p, s->col_idx.p, nmesh, node_off, cost);
}

static int batch_prepare(pf_system *s) {
  if (!s->batch) {
    pf_set_error("batch: call pf_system_batch_set first");
    return PF_INVALID;
  }
  return apply_bc(s);
}

extern "C" int pf_system_batch_solve(pf_system *s, const pf_solve_opts *opts, const pf_batch_tune *tune,
                                     double *x_out, pf_batch_mesh_stats *mesh_stats, pf_stats *total) {
  PF_CHECK(sys_check(s, true));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  pf_solve_opts o;
  if (opts) o = *opts; else pf_solve_opts_default(&o);
  if (o.preconditioning != 0 && o.preconditioning != 1) {
    pf_set_error("batch: preconditioning must be 0 (none) or 1 (PointJacobi)");
    return PF_INVALID;
  }
  PF_CHECK(batch_prepare(s));
  if (o.preconditioning == 1) {  // JacobiPreconditionerCRS is row-local: the union's = every mesh's
    PF_CHECK(pf_jacobi_rows(ctx, s->n, s->row_ptr.p, s->col_idx.p, s->val.p, s->rhs.p));
    mark_changed(s);
    ++s->rhs_version;
  }
  const int rc = pf_batch_solve_device(ctx, s->batch, s->row_ptr.p, s->col_idx.p, s->val.p, s->nnz, s->rhs.p,
                                       s->x.p, o, tune, s->mat_version, mesh_stats, total);
  if (rc < 0 || rc == PF_INVALID) return rc;
  if (s->has_fix) PF_CHECK(pf_bc_set_fixed_x(ctx, s->n, s->fixed.p, s->ubar.p, s->x.p));
  if (x_out) PF_CHECK(download(ctx, x_out, s->x.p, (size_t)s->n));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return rc;
}

extern "C" int pf_system_batch_autotune(pf_system *s, const pf_solve_opts *opts, int32_t iters, pf_batch_tune *best,
                                        int32_t max_results, pf_batch_tune *tried, double *tried_ms,
                                        int32_t *ntried) {
  PF_CHECK(sys_check(s, true));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  pf_solve_opts o;
  if (opts) o = *opts; else pf_solve_opts_default(&o);
  PF_CHECK(batch_prepare(s));
  return pf_batch_autotune_device(ctx, s->batch, s->row_ptr.p, s->col_idx.p, s->val.p, s->nnz, s->rhs.p, s->x.p, o,
                                  iters, s->mat_version, best, max_results, tried, tried_ms, ntried);
}

// ============================================================================ multi-RHS / adjoint
// ---- direct solve: solve(algorithm="GaussJordan") (FEMSolverClass.f90:1795-1798) ----------

static int direct_finish(pf_system *s, int rc, double *x_out) {
  pf_ctx *ctx = s->ctx;
  if (rc < 0 || rc == PF_INVALID) return rc;
  if (s->has_fix) PF_CHECK(pf_bc_set_fixed_x(ctx, s->n, s->fixed.p, s->ubar.p, s->x.p));
  if (x_out) PF_CHECK(download(ctx, x_out, s->x.p, (size_t)s->n));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return rc;
}

static const double *direct_coords(pf_system *s) {
  return (s->g.has_coords && 3 * s->g.nn == s->n) ? s->g.nodcoord.p : nullptr;
}

extern "C" int pf_system_solve_direct(pf_system *s, int32_t refine, double *x_out, pf_stats *stats) {
  PF_CHECK(sys_check(s, true));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  PF_CHECK(apply_bc(s));
  const int64_t off[2] = {0, s->n};
  PF_CHECK(pf_direct_analyse(ctx, &s->direct, s->n, s->row_ptr.p, s->col_idx.p, 1, off, direct_coords(s), s->g.nn,
                             s->pattern_version));
  pf_stats st;
  const int rc = pf_direct_solve_device(ctx, s->direct, s->n, s->row_ptr.p, s->col_idx.p, s->val.p, s->nnz,
                                        s->rhs.p, s->x.p, s->mat_version, refine, s->direct_mfma, nullptr, &st);
  if (stats && rc >= 0 && rc != PF_INVALID) *stats = st;
  return direct_finish(s, rc, x_out);
}

extern "C" int pf_system_batch_direct(pf_system *s, int32_t refine, double *x_out, pf_batch_mesh_stats *mesh_stats,
                                      pf_stats *total) {
  PF_CHECK(sys_check(s, true));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  PF_CHECK(batch_prepare(s));
  std::vector<int64_t> off;
  PF_CHECK(pf_batch_dof_offsets(s->batch, off));
  const int32_t nmesh = (int32_t)off.size() - 1;
  PF_CHECK(pf_direct_analyse(ctx, &s->direct, s->n, s->row_ptr.p, s->col_idx.p, nmesh, off.data(), direct_coords(s),
                             s->g.nn, s->pattern_version));
  pf_stats st;
  const int rc = pf_direct_solve_device(ctx, s->direct, s->n, s->row_ptr.p, s->col_idx.p, s->val.p, s->nnz,
                                        s->rhs.p, s->x.p, s->mat_version, refine, s->direct_mfma, mesh_stats, &st);
  if (total && rc >= 0 && rc != PF_INVALID) *total = st;
  return direct_finish(s, rc, x_out);
}

extern "C" int pf_system_direct_info(pf_system *s, int32_t mesh, int64_t *n, int32_t *bandwidth, int32_t *ordering,
                                     int32_t *perm, double *factor) {
  PF_CHECK(sys_check(s, false));
  DeviceGuard guard(s->ctx);
  PF_CHECK(pf_direct_info(s->direct, mesh, n, bandwidth, ordering, perm));
  if (factor) PF_CHECK(pf_direct_factor_get(s->ctx, s->direct, mesh, factor));
  return PF_OK;
}

extern "C" int pf_system_direct_kernel(pf_system *s, int32_t mfma) {
  PF_CHECK(sys_check(s, false));
  s->direct_mfma = mfma ? 1 : 0;
  pf_direct_invalidate(s->direct);
  return PF_OK;
}

extern "C" int pf_system_ke_kernel(pf_system *s, int32_t kind) {
  PF_CHECK(sys_check(s, false));
  if (kind < 0 || kind > 2) {
    pf_set_error("pf_system_ke_kernel: kind must be 0 (reference order), 1 (isotropic split, VALU) or 2 "
                 "(isotropic split, MFMA)");
    return PF_INVALID;
  }
  s->ke_kernel = kind;
  return PF_OK;
}

extern "C" int pf_system_solve_multi_dev(pf_system *s, int32_t nrhs, const double *B_dev, double *X_dev,
                                         const pf_solve_opts *opts, pf_stats *per_rhs, pf_stats *total) {
  PF_CHECK(sys_check(s, true));
  if (!B_dev || !X_dev) {
    pf_set_error("pf_system_solve_multi_dev: device arrays B and X required");
    return PF_INVALID;
  }
  if (s->ubar_nonzero) {
    pf_set_error("pf_system_solve_multi_dev: nonzero prescribed displacements are not supported (fix with 0)");
    return PF_INVALID;
  }
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  pf_solve_opts o;
  if (opts) o = *opts; else pf_solve_opts_default(&o);
  PF_CHECK(apply_bc(s));
  return pf_multi_solve_device(ctx, &s->multi, s->n, s->row_ptr.p, s->col_idx.p, s->val.p, s->mat_version,
                               s->has_fix ? s->fixed.p : nullptr, nrhs, B_dev, X_dev, o, per_rhs, total);
}

extern "C" int pf_system_elem_sensitivity_dev(pf_system *s, int32_t nrhs, const double *U_dev, const double *Lam_dev,
                                              double *out_dev) {
  PF_CHECK(sys_check(s, true));
  if (!s->has_mat || s->dof != 3 || nrhs < 1 || !U_dev || !Lam_dev || !out_dev) {
    pf_set_error("pf_system_elem_sensitivity_dev: assembled hex8 stiffness, nrhs >= 1 and device arrays required");
    return PF_INVALID;
  }
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  PF_CHECK(pf_elem_bilinear(ctx, s->g, s->E.p, s->nu.p, s->jmode, nrhs, U_dev, Lam_dev,
                            s->has_fix ? s->fixed.p : nullptr, out_dev));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

// ============================================================================ body force / stress
extern "C" int pf_system_add_body_force(pf_system *s, const double *density, double density_all,
                                        const double *accel) {
  PF_CHECK(sys_check(s, true));
  if (!s->has_mesh || s->g.nne != 8 || s->dof != 3 || !accel) {
    pf_set_error("pf_system_add_body_force: hex8 mesh with a DOF=3 CRS and an acceleration required");
    return PF_INVALID;
  }
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  if (!s->g.built) PF_CHECK(pf_graph_node_elems(ctx, s->g));
  DevBuf<double> drho;
  if (density) PF_CHECK(upload(ctx, drho, density, (size_t)s->g.ne));
  PF_CHECK(pf_body_force(ctx, s->g, density ? drho.p : nullptr, density_all, accel, s->rhs.p));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  ++s->rhs_version;
  return PF_OK;
}

extern "C" int pf_system_element_stress(pf_system *s, const double *E, int64_t nE, const double *nu, int64_t nnu,
                                        int32_t jmode, double *sigma) {
  PF_CHECK(sys_check(s, true));
  if (!s->has_mesh || s->g.nne != 8 || s->dof != 3 || !sigma) {
    pf_set_error("pf_system_element_stress: hex8 mesh with a DOF=3 system required");
    return PF_INVALID;
  }
  PF_CHECK(stress_params(s->g.ne, nE, nnu, E, nu));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  DevBuf<double> dE, dnu, ds;
  PF_CHECK(upload(ctx, dE, E, (size_t)nE));
  PF_CHECK(upload(ctx, dnu, nu, (size_t)nnu));
  PF_CHECK(ds.ensure((size_t)s->g.ne * 9 + 1));
  PF_CHECK(pf_cell_stress(ctx, s->g, s->x.p, dE.p, nE == s->g.ne && s->g.ne > 1 ? 1 : 0, dnu.p,
                          nnu == s->g.ne && s->g.ne > 1 ? 1 : 0, jmode, ds.p));
  return download(ctx, sigma, ds.p, (size_t)s->g.ne * 9);
}

// ====================================================== multi-domain assembly with overset
static int check_hex8_mesh_args(int64_t nn, int64_t ne, const double *xyz, const int32_t *en) {
  if (nn < 0 || ne < 0 || (nn > 0 && !xyz) || (ne > 0 && !en) || nn >= (int64_t)INT32_MAX ||
      ne * 8 >= (int64_t)INT32_MAX) {
    pf_set_error("invalid hex8 mesh arguments");
    return PF_INVALID;
  }
  return PF_OK;
}

static int overset_common(pf_ctx *ctx, bool p2p, int64_t nnA, int64_t neA, const double *nodcoordA,
                          const int32_t *elemnodA, int64_t nnB, int64_t neB, const double *nodcoordB,
                          const int32_t *elemnodB, int64_t cap, int64_t *count, int32_t *own, int32_t *gauss_point,
                          int32_t *pair_elem, double *position) {
  PF_CHECK(check_ctx(ctx));
  PF_CHECK(check_hex8_mesh_args(nnA, neA, nodcoordA, elemnodA));
  PF_CHECK(check_hex8_mesh_args(nnB, neB, nodcoordB, elemnodB));
  if (!count || cap < 0 || (cap > 0 && (!own || (!p2p && !gauss_point) || !pair_elem || !position))) {
    pf_set_error("overset: count and (for cap > 0) the output arrays required");
    return PF_INVALID;
  }
  DeviceGuard guard(ctx);
  NodeGraph A, B;
  PF_CHECK(set_mesh(ctx, A, nnA, neA, 8, nodcoordA, elemnodA));
  PF_CHECK(pf_graph_check_mesh(ctx, A));
  PF_CHECK(set_mesh(ctx, B, nnB, neB, 8, nodcoordB, elemnodB));
  PF_CHECK(pf_graph_node_elems(ctx, B));
  DevBuf<int32_t> e, g, p;
  DevBuf<double> x;
  int64_t k = 0;
  PF_CHECK(pf_overset_search(ctx, A, B, p2p, e, g, p, x, &k));
  *count = k;
  const int64_t m = std::min(k, cap);
  if (m > 0) {
    PF_CHECK(download(ctx, own, e.p, (size_t)m));
    if (!p2p) PF_CHECK(download(ctx, gauss_point, g.p, (size_t)m));
    PF_CHECK(download(ctx, pair_elem, p.p, (size_t)m));
    PF_CHECK(download(ctx, position, x.p, (size_t)m * 3));
  }
  return PF_OK;
}

extern "C" int pf_overset_gpp(pf_ctx *ctx, int64_t nnA, int64_t neA, const double *nodcoordA,
                              const int32_t *elemnodA, int64_t nnB, int64_t neB, const double *nodcoordB,
                              const int32_t *elemnodB, int64_t cap, int64_t *count, int32_t *elem,
                              int32_t *gauss_point, int32_t *pair_elem, double *position) {
  return overset_common(ctx, false, nnA, neA, nodcoordA, elemnodA, nnB, neB, nodcoordB, elemnodB, cap, count, elem,
                        gauss_point, pair_elem, position);
}

extern "C" int pf_overset_p2p(pf_ctx *ctx, int64_t nnA, int64_t neA, const double *nodcoordA,
                              const int32_t *elemnodA, int64_t nnB, int64_t neB, const double *nodcoordB,
                              const int32_t *elemnodB, int64_t cap, int64_t *count, int32_t *node,
                              int32_t *pair_elem, double *position) {
  return overset_common(ctx, true, nnA, neA, nodcoordA, elemnodA, nnB, neB, nodcoordB, elemnodB, cap, count, node,
                        nullptr, pair_elem, position);
}

// own: element id (GPP) or node id (P2P, Gauss point id 0) of the joint mesh, 1-based
static int check_connects(const pf_system *s, int64_t ncon, const int32_t *elem, const int32_t *gauss_point,
                          const int32_t *pair_elem) {
  for (int64_t c = 0; c < ncon; ++c) {
    const bool p2p = gauss_point && gauss_point[c] == 0;
    const int64_t lim = p2p ? s->g.nn : s->g.ne;
    if (elem[c] < 1 || elem[c] > lim || pair_elem[c] < 1 || pair_elem[c] > s->g.ne ||
        (gauss_point && (gauss_point[c] < 0 || gauss_point[c] > 8))) {
      pf_set_error("connect %lld: own %s id, paired element id or Gauss point id out of range", (long long)c + 1,
                   p2p ? "node" : "element");
      return PF_INVALID;
    }
  }
  return PF_OK;
}

extern "C" int pf_system_multi_crs(pf_system *s, int32_t dof, int32_t any_overset, int64_t ncon,
                                   const int32_t *elem, const int32_t *gauss_point, const int32_t *pair_elem,
                                   int64_t *nnz) {
  PF_CHECK(sys_check(s, false));
  if (!s->has_mesh || s->g.nne != 8 || dof < 1 || ncon < 0 || (ncon > 0 && (!elem || !pair_elem))) {
    pf_set_error("pf_system_multi_crs: hex8 mesh, dof >= 1 and the connect arrays required");
    return PF_INVALID;
  }
  PF_CHECK(check_connects(s, ncon, elem, gauss_point, pair_elem));
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  s->dof = dof;
  s->n = (int64_t)dof * s->g.nn;
  PF_CHECK(s->row_ptr.ensure((size_t)s->n + 1));
  if (!any_overset) {  // no domain holds overset connects: every row stays empty
    std::vector<int64_t> ones((size_t)s->n + 1, 1);
    PF_CHECK(upload(ctx, s->row_ptr, ones.data(), ones.size()));
    s->nnz = 0;
    PF_CHECK(s->col_idx.ensure(1));
    PF_CHECK(s->val.ensure(1));
  } else {
    DevBuf<int32_t> dmy, dpr, dgp;
    if (ncon > 0) {
      std::vector<int32_t> a((size_t)ncon), b((size_t)ncon);
      for (int64_t c = 0; c < ncon; ++c) {
        a[c] = elem[c] - 1;
        b[c] = pair_elem[c] - 1;
      }
      PF_CHECK(upload(ctx, dmy, a.data(), a.size()));
      PF_CHECK(upload(ctx, dpr, b.data(), b.size()));
      if (gauss_point) PF_CHECK(upload(ctx, dgp, gauss_point, (size_t)ncon));
    }
    NodeGraph merged;
    PF_CHECK(pf_graph_merged_cols(ctx, s->g, ncon, dmy.p, dpr.p, gauss_point ? dgp.p : nullptr, merged));
    s->nnz = nnz_of(merged, dof);
    if (s->nnz >= (int64_t)1 << 40) {
      pf_set_error("CRS too large");
      return PF_INVALID;
    }
    PF_CHECK(s->col_idx.ensure((size_t)s->nnz + 1));
    PF_CHECK(s->val.ensure((size_t)s->nnz + 1));
    PF_CHECK(pf_graph_expand(ctx, merged, dof, s->row_ptr.p, s->col_idx.p));
  }
  PF_HIP(hipMemsetAsync(s->val.p, 0, sizeof(double) * (s->nnz + 1), ctx->stream));
  PF_CHECK(sys_alloc_vectors(s, true));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  s->graph_crs = false;
  s->multi_crs = true;
  ++s->pattern_version;
  mark_changed(s);
  if (nnz) *nnz = s->nnz;
  return PF_OK;
}

extern "C" int pf_system_ebom(pf_system *s, int64_t ncon, const int32_t *elem, const int32_t *gauss_point,
                              const int32_t *pair_elem, const double *position, double penalty, int32_t dof) {
  PF_CHECK(sys_check(s, true));
  if (!s->has_mesh || s->g.nne != 8 || !s->g.has_coords || ncon < 0 ||
      (ncon > 0 && (!elem || !gauss_point || !pair_elem || !position))) {
    pf_set_error("pf_system_ebom: hex8 mesh with coordinates and the connect arrays required");
    return PF_INVALID;
  }
  if (dof != 3 || s->dof != 3) {
    pf_set_error("pf_system_ebom: DOF = 3 (3-D hex8 displacement) is the supported form");
    return PF_INVALID;
  }
  PF_CHECK(check_connects(s, ncon, elem, gauss_point, pair_elem));
  if (ncon == 0) return PF_OK;
  pf_ctx *ctx = s->ctx;
  DeviceGuard guard(ctx);
  std::vector<int32_t> a((size_t)ncon), b((size_t)ncon);
  for (int64_t c = 0; c < ncon; ++c) {
    a[c] = elem[c] - 1;
    b[c] = pair_elem[c] - 1;
  }
  DevBuf<int32_t> dmy, dgp, dpr, err;
  DevBuf<double> dpos;
  PF_CHECK(upload(ctx, dmy, a.data(), a.size()));
  PF_CHECK(upload(ctx, dpr, b.data(), b.size()));
  PF_CHECK(upload(ctx, dgp, gauss_point, (size_t)ncon));
  PF_CHECK(upload(ctx, dpos, position, (size_t)ncon * 3));
  PF_CHECK(err.ensure(1));
  PF_HIP(hipMemsetAsync(err.p, 0, sizeof(int32_t), ctx->stream));
  PF_CHECK(pf_ebom_assemble(ctx, s->g, ncon, dmy.p, dgp.p, dpr.p, dpos.p, penalty, s->row_ptr.p, s->col_idx.p,
                            s->multi_crs, s->val.p, err.p));
  int32_t herr = 0;
  PF_CHECK(download(ctx, &herr, err.p, 1));
  mark_changed(s);
  if (herr) {
    pf_set_error("setEbOM: no CRS slot for a connect entry (setCRS before the overset?)");
    return PF_INVALID;
  }
  return PF_OK;
}
