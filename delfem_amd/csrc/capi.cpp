// C-ABI of libdfm_gpu.so (include/dfm_gpu.h): host orchestration of the device system.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>
#include <unordered_map>
#include <mutex>
#include <map>
#include <stdexcept>

#include "dfm_internal.hpp"
#include "kernels.hpp"
#include "solver.hpp"

namespace dfm {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

static std::atomic<long long> g_dev_bytes{0};  // bytes handed out (cached blocks not counted)

namespace {
constexpr int kMaxDevices = 64;
constexpr size_t kCacheCap = 16ull << 30;  // cached device bytes per device
constexpr size_t kHostCacheCap = 64ull << 20;
thread_local int g_cache_guard = 0;

size_t size_class(size_t b) {
  if (b <= 4096) return (b + 255) & ~size_t(255);
  int k = 63 - __builtin_clzll((unsigned long long)b);
  const size_t step = size_t(1) << (k - 3);
  return (b + step - 1) & ~(step - 1);
}

struct Pool {
  std::mutex mu;
  std::multimap<size_t, void*> free_blocks[kMaxDevices + 1];  // [kMaxDevices]: pinned host memory
  std::unordered_map<void*, std::pair<int, size_t>> owner;    // block -> (device or kMaxDevices, class size)
  size_t cached[kMaxDevices + 1] = {};
};
Pool& pool() {
  static Pool* p = new Pool;  // never destroyed: releases may come from static destructors
  return *p;
}

hipError_t pool_alloc(void** p, size_t bytes, bool host) {
  const size_t cls = size_class(std::max<size_t>(bytes, 1));
  int dev = kMaxDevices;
  if (!host) {
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = 0;
  }
  Pool& P = pool();
  {
    std::lock_guard<std::mutex> lk(P.mu);
    auto& fb = P.free_blocks[dev];
    auto it = fb.find(cls);
    if (it != fb.end()) {
      *p = it->second;
      fb.erase(it);
      P.cached[dev] -= cls;
      if (!host) g_dev_bytes += (long long)cls;
      return hipSuccess;
    }
  }
  hipError_t e = host ? hipHostMalloc(p, cls) : hipMalloc(p, cls);
  if (e != hipSuccess && !host) {
    // out of memory with blocks cached: return them and retry once
    std::vector<void*> drop;
    {
      std::lock_guard<std::mutex> lk(P.mu);
      for (auto& kv : P.free_blocks[dev]) drop.push_back(kv.second);
      P.free_blocks[dev].clear();
      P.cached[dev] = 0;
      for (void* q : drop) P.owner.erase(q);
    }
    for (void* q : drop) (void)hipFree(q);
    (void)hipGetLastError();
    e = hipMalloc(p, cls);
  }
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(P.mu);
  P.owner[*p] = {dev, cls};
  if (!host) g_dev_bytes += (long long)cls;
  return hipSuccess;
}

void pool_release(void* p, bool host) {
  Pool& P = pool();
  std::vector<void*> drop;
  int dev = 0;
  {
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.owner.find(p);
    if (it == P.owner.end()) return;
    dev = it->second.first;
    const size_t cls = it->second.second;
    if (!host) g_dev_bytes -= (long long)cls;
    if (g_cache_guard == 0) {
      P.owner.erase(it);
      drop.push_back(p);
    } else {
      P.free_blocks[dev].insert({cls, p});
      P.cached[dev] += cls;
      const size_t cap = host ? kHostCacheCap : kCacheCap;
      auto& fb = P.free_blocks[dev];
      while (P.cached[dev] > cap && !fb.empty()) {
        auto last = std::prev(fb.end());
        P.cached[dev] -= last->first;
        P.owner.erase(last->second);
        drop.push_back(last->second);
        fb.erase(last);
      }
    }
  }
  for (void* q : drop) {
    if (host) (void)hipHostFree(q);
    else (void)hipFree(q);
  }
}
}  // namespace

hipError_t DevAlloc::alloc(void** p, size_t bytes) { return pool_alloc(p, bytes, false); }
void DevAlloc::release(void* p) { pool_release(p, false); }
hipError_t DevAlloc::alloc_host(void** p, size_t bytes) { return pool_alloc(p, bytes, true); }
void DevAlloc::release_host(void* p) { pool_release(p, true); }
CacheRelease::CacheRelease() { g_cache_guard++; }
CacheRelease::~CacheRelease() { g_cache_guard--; }

long long device_bytes_in_use() { return g_dev_bytes.load(); }

Sys::~Sys() {
  // no queued work on the system's stream may touch the blocks once they are cached for reuse (dfm_sys_destroy
  // has synchronised it already; failure paths of the create functions may not have)
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  CacheRelease guard;
  d_rowptr.release(); d_colind.release(); d_diaind.release(); d_lsplit.release(); d_usplit.release();
  d_frowptr.release(); d_fcolind.release(); d_fdiaind.release(); d_a2f.release();
  d_order.release(); d_inv.release(); d_elsup_ptr.release(); d_elsup.release(); d_lnods.release(); d_bc.release();
  d_eptr.release(); d_eidx.release(); d_eloc.release(); d_rowptr0.release(); d_colind0.release(); d_permq.release();
  d_elnod.release();
  d_coords.release(); d_u.release(); d_v.release(); d_a.release(); d_v0.release();
  d_crs.release(); d_dia.release(); d_crs32.release(); d_dia32.release();
  d_fcrs.release(); d_fdia.release();
  d_res.release(); d_upd.release(); d_rhs.release();
  for (auto* t : d_tmp) { t->release(); delete t; }
  d_work.release(); d_work32.release(); d_partial.release(); d_part_sw.release(); d_counter.release(); d_scal.release(); d_iflag.release();
  for (Sweep* w : {&fwd, &bwd}) w->release_dev();
  sell.release_dev();
  fplan.release_dev();
  d_par.release(); d_lptr.release(); d_loads.release();
  d_force.release(); d_bptr.release(); d_bcol.release(); d_bsrc.release(); d_bval.release();
  for (auto& g : gexec)
    if (g) (void)hipGraphExecDestroy(g);
  delete last_asm;
  if (h_pinned) DevAlloc::release_host(h_pinned);
  if (h_pinned_d) DevAlloc::release_host(h_pinned_d);
  for (hipEvent_t ev : poll_ev)
    if (ev) (void)hipEventDestroy(ev);
}

template <typename T>
static hipError_t upload(DBuf<T>& b, const std::vector<T>& v, hipStream_t st) {
  hipError_t e = b.alloc(v.size());
  if (e != hipSuccess || v.empty()) return e;
  e = hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

// device pattern, ordering maps and sweep schedules for the current order
// Factor pattern and sweep schedule (device copies); the slab values are filled by dfm_prec_factor.
static int upload_sweeps(Sys& s) {
  hipStream_t st = s.ctx->stream;
  // the launch shapes of a captured Krylov chunk (sweep block counts, phase ranges, SELL slices, register caps)
  // follow this schedule: a new one invalidates the graphs even where the allocator hands back the same addresses
  s.struct_gen++;
  DFM_HIP(upload(s.d_lsplit, s.lsplit, st));
  DFM_HIP(upload(s.d_usplit, s.usplit, st));
  DFM_HIP(upload(s.d_frowptr, s.frowptr, st));
  DFM_HIP(upload(s.d_fcolind, s.fcolind, st));
  DFM_HIP(upload(s.d_fdiaind, s.fdiaind, st));
  DFM_HIP(upload(s.d_a2f, s.a2f, st));
  for (Sweep* w : {&s.fwd, &s.bwd}) {
    w->release_dev();
    DFM_HIP(upload(w->d_grp, w->grp, st));
    DFM_HIP(upload(w->d_rowlist, w->rowlist, st));
    DFM_HIP(upload(w->d_lev, w->lev, st));
    DFM_HIP(upload(w->d_gmeta, w->gmeta, st));
    DFM_HIP(upload(w->d_srow, w->srow, st));
    DFM_HIP(upload(w->d_scnt, w->scnt, st));
    DFM_HIP(upload(w->d_ecol, w->ecol, st));
    DFM_HIP(upload(w->d_esrc, w->esrc, st));
    DFM_HIP(upload(w->d_fuse_bgrp, w->fuse_bgrp, st));
  }
  FactorPlan& fp = s.fplan;
  fp.release_dev();
  if (fp.ok) {
    DFM_HIP(upload(fp.d_gpass, fp.gpass, st));
    DFM_HIP(upload(fp.d_pmeta, fp.pmeta, st));
    DFM_HIP(upload(fp.d_etab, fp.etab, st));
    DFM_HIP(upload(fp.d_utab, fp.utab, st));
    DFM_HIP(upload(fp.d_ukj, fp.ukj, st));
    DFM_HIP(upload(fp.d_ustep, fp.ustep, st));
  }
  return DFM_OK;
}

// Device copies of the matrix pattern, its SELL layout (values filled at assembly) and the sweeps.
static int upload_structure(Sys& s) {
  hipStream_t st = s.ctx->stream;
  DFM_HIP(upload(s.d_rowptr, s.rowptr, st));
  DFM_HIP(upload(s.d_colind, s.colind, st));
  DFM_HIP(upload(s.d_diaind, s.diaind, st));
  DFM_HIP(upload(s.d_order, s.order, st));
  DFM_HIP(upload(s.d_inv, s.inv, st));
  DFM_HIP(upload(s.d_eptr, s.eptr, st));
  DFM_HIP(upload(s.d_eidx, s.eidx, st));
  if (s.perm_crs.size() < 0xFFFFFFFFull) {
    std::vector<unsigned> pq(s.perm_crs.begin(), s.perm_crs.end());
    DFM_HIP(upload(s.d_permq, pq, st));  // (upload waits for the copy)
  }
  s.sell.release_dev();
  DFM_HIP(upload(s.sell.d_sptr, s.sell.sptr, st));
  DFM_HIP(upload(s.sell.d_visit, s.sell.visit, st));
  DFM_HIP(upload(s.sell.d_col, s.sell.col, st));
  DFM_HIP(upload(s.sell.d_src, s.sell.src, st));
  return upload_sweeps(s);
}

// Dirichlet flags into the device numbering (again after a new ordering)
static int upload_bc(Sys& s) {
  if (s.bc0.empty()) return DFM_OK;
  std::vector<int> bc((size_t)s.n * s.len);
  for (int i0 = 0; i0 < s.n; i0++)
    for (int d = 0; d < s.len; d++) bc[(size_t)i0 * s.len + d] = s.bc0[(size_t)s.order[i0] * s.len + d];
  DFM_HIP(hipMemcpy(s.d_bc.p, bc.data(), bc.size() * sizeof(int), hipMemcpyHostToDevice));
  return DFM_OK;
}

static int apply_order(Sys& s, const std::vector<int>& order, const std::vector<int>& color,
                       const std::vector<int>& tile, int ncolors) {
  s.order = order;
  s.inv.assign(s.n, 0);
  for (int i = 0; i < s.n; i++) s.inv[order[i]] = i;
  s.natural = true;
  for (int i = 0; i < s.n; i++)
    if (order[i] != i) { s.natural = false; break; }
  s.color_of = color;
  s.tile_of = tile;
  s.ncolors = ncolors;
  try {
    build_device_pattern(s);
    build_schedule(s);
  } catch (const std::exception& ex) {
    set_error(ex.what());
    return DFM_ERR_ARG;
  }
  s.assembled = s.factored = s.factored32 = false;
  s.saved = SAVE_NONE;
  const int rc = upload_structure(s);
  if (rc != DFM_OK) return rc;
  return upload_bc(s);
}

static double* vec_ptr(Sys& s, int iv) {
  if (iv == -1) return s.d_res.p;
  if (iv == -2) return s.d_upd.p;
  if (iv >= 0 && iv < (int)s.d_tmp.size()) return s.d_tmp[iv]->p;
  return nullptr;
}

static long long mdof(const Sys& s) { return (long long)s.n * s.len; }

double* sys_vec_ptr(Sys& s, int iv) { return vec_ptr(s, iv); }

}  // namespace dfm

using namespace dfm;

#define SYS(h) (h ? &h->s : nullptr)
#define NEED(p, name) DFM_CHECK((p) != nullptr, DFM_ERR_ARG, std::string(name) + ": null argument")

extern "C" {

const char* dfm_last_error(void) { return g_err.c_str(); }

int dfm_device_count(int* count) {
  NEED(count, "dfm_device_count");
  int c = 0;
  DFM_HIP(hipGetDeviceCount(&c));
  *count = c;
  return DFM_OK;
}

int dfm_ctx_create(int device, dfm_ctx* out) {
  NEED(out, "dfm_ctx_create");
  DFM_HIP(hipSetDevice(device));
  dfm_ctx c = new dfm_ctx_s;
  c->c.device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    set_error(std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
    return DFM_ERR_HIP;
  }
  *out = c;
  return DFM_OK;
}

}  // extern "C"

void ctx_release(dfm_ctx ctx) {
  if (ctx->c.refs.fetch_sub(1) != 1) return;
  (void)hipStreamSynchronize(ctx->c.stream);
  {
    CacheRelease guard;  // the stream is idle: the buffer can go back to the cache
    ctx->c.ebuf.release();
  }
  (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

extern "C" {

// Systems keep their context alive: destroying a context with live systems defers the release.
int dfm_ctx_destroy(dfm_ctx ctx) {
  if (!ctx || ctx->c.destroyed) return DFM_OK;
  ctx->c.destroyed = true;
  ctx_release(ctx);
  return DFM_OK;
}

int dfm_ctx_synchronize(dfm_ctx ctx) {
  NEED(ctx, "dfm_ctx_synchronize");
  DFM_HIP(hipStreamSynchronize(ctx->c.stream));
  return DFM_OK;
}

// Device copies of the mesh, value arrays, vectors and scalars of a system whose host structure is set.
static hipError_t sys_alloc_device(Sys& s) {
  hipStream_t st = s.ctx->stream;
  const long long m = mdof(s);
  const int nnode = s.n;
  hipError_t e;
  if ((e = upload(s.d_coords, s.coords, st)) != hipSuccess) return e;
  if ((e = upload(s.d_lnods, s.lnods, st)) != hipSuccess) return e;
  if ((e = upload(s.d_elsup_ptr, s.elsup_ptr, st)) != hipSuccess) return e;
  if ((e = upload(s.d_elsup, s.elsup, st)) != hipSuccess) return e;
  if ((e = upload(s.d_eloc, s.eloc, st)) != hipSuccess) return e;
  if ((e = upload(s.d_rowptr0, s.rowptr0, st)) != hipSuccess) return e;
  if ((e = upload(s.d_colind0, s.colind0, st)) != hipSuccess) return e;
  // two doubles of padding: the factor kernel reads 3x3 blocks as five aligned 16-byte loads, one past the end
  if ((e = s.d_crs.alloc((size_t)s.ncrs * s.bs + 2)) != hipSuccess) return e;
  if ((e = s.d_dia.alloc((size_t)nnode * s.bs + 2)) != hipSuccess) return e;
  if ((e = s.d_res.alloc(m)) != hipSuccess) return e;
  if ((e = s.d_upd.alloc(m)) != hipSuccess) return e;
  if ((e = s.d_rhs.alloc(m)) != hipSuccess) return e;
  if ((e = s.d_u.alloc(m)) != hipSuccess) return e;
  if ((e = s.d_v.alloc(m)) != hipSuccess) return e;
  if ((e = s.d_bc.alloc(m)) != hipSuccess) return e;
  // per-block partial sums: up to 2 values per block of the widest reduction grid (SpMV rows at 32 lanes)
  if ((e = s.d_partial.alloc(8 * 2048)) != hipSuccess) return e;
  if ((e = s.d_counter.alloc(8)) != hipSuccess) return e;
  if ((e = s.d_scal.alloc(kNScal)) != hipSuccess) return e;
  if ((e = s.d_iflag.alloc(8)) != hipSuccess) return e;
  if ((e = DevAlloc::alloc_host((void**)&s.h_pinned, 2 * 16 * sizeof(int))) != hipSuccess) return e;
  if ((e = DevAlloc::alloc_host((void**)&s.h_pinned_d, 2 * kNScal * sizeof(double))) != hipSuccess) return e;
  if ((e = hipMemsetAsync(s.d_counter.p, 0, 8 * sizeof(unsigned), st)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(s.d_bc.p, 0, m * sizeof(int), st)) != hipSuccess) return e;
  if ((e = launch_fill(s.d_u.p, 0.0, m, st)) != hipSuccess) return e;
  if ((e = launch_fill(s.d_v.p, 0.0, m, st)) != hipSuccess) return e;
  if ((e = launch_fill(s.d_res.p, 0.0, m, st)) != hipSuccess) return e;
  if ((e = launch_fill(s.d_upd.p, 0.0, m, st)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  return hipSuccess;
}

int dfm_sys_create_mesh(dfm_ctx ctx, int ndim, int nnode, const double* coords, int nelem, const int* lnods, int len,
                        dfm_sys* out) {
  NEED(ctx, "dfm_sys_create_mesh");
  NEED(out, "dfm_sys_create_mesh");
  DFM_CHECK(ndim == 2 || ndim == 3, DFM_ERR_ARG, "dfm_sys_create_mesh: ndim must be 2 or 3");
  DFM_CHECK(len >= 1 && len <= 3, DFM_ERR_ARG, "dfm_sys_create_mesh: len must be 1, 2 or 3");
  DFM_CHECK(nnode > 0 && nelem > 0 && coords && lnods, DFM_ERR_ARG, "dfm_sys_create_mesh: empty mesh");
  const int npoel = ndim + 1;
  for (long long k = 0; k < (long long)nelem * npoel; k++)
    DFM_CHECK(lnods[k] >= 0 && lnods[k] < nnode, DFM_ERR_ARG, "dfm_sys_create_mesh: node index out of range");
  dfm_sys h = new dfm_sys_s;
  Sys& s = h->s;
  s.ctx = &ctx->c;
  s.ndim = ndim; s.n = nnode; s.len = len; s.bs = len * len; s.npoel = npoel; s.nelem = nelem;
  s.coords.assign(coords, coords + (size_t)nnode * ndim);
  s.lnods.assign(lnods, lnods + (size_t)nelem * npoel);
  build_pattern(nnode, nelem, npoel, lnods, s.rowptr0, s.colind0, s.elsup_ptr, s.elsup);
  s.ncrs = s.rowptr0[nnode];
  std::vector<int> ident(nnode);
  std::iota(ident.begin(), ident.end(), 0);
  int rc = apply_order(s, ident, {}, {}, 0);
  if (rc != DFM_OK) { delete h; return rc; }
  auto fail = [&](hipError_t e) {
    set_error(std::string("dfm_sys_create_mesh: ") + hipGetErrorString(e));
    delete h;
    return e == hipErrorOutOfMemory ? DFM_ERR_NOMEM : DFM_ERR_HIP;
  };
  const hipError_t e = sys_alloc_device(s);
  if (e != hipSuccess) return fail(e);
  h->owner = ctx;
  ctx->c.refs++;
  *out = h;
  return DFM_OK;
}

// CMatDia_BlkCrs::AddPattern(const CIndexedArray&) (src/matvec/matdia_blkcrs.cpp:346-505) on an empty matrix: the
// block pattern as given (off-diagonal blocks only; rows past the pattern's are empty), no mesh behind it.
int dfm_sys_create_pattern(dfm_ctx ctx, int nblk, int len, int nrow, const unsigned* rowptr, const unsigned* colind,
                           dfm_sys* out) {
  NEED(ctx, "dfm_sys_create_pattern");
  NEED(out, "dfm_sys_create_pattern");
  DFM_CHECK(len >= 1 && len <= 3, DFM_ERR_ARG, "dfm_sys_create_pattern: len must be 1, 2 or 3");
  DFM_CHECK(nblk > 0 && nrow >= 0 && nrow <= nblk, DFM_ERR_ARG, "dfm_sys_create_pattern: 0 <= nrow <= nblk > 0");
  DFM_CHECK(nrow == 0 || (rowptr && colind), DFM_ERR_ARG, "dfm_sys_create_pattern: null pattern");
  DFM_CHECK(nrow == 0 || rowptr[0] == 0, DFM_ERR_ARG, "dfm_sys_create_pattern: rowptr[0] must be 0");
  std::vector<unsigned> rp((size_t)nblk + 1, 0), ci;
  std::vector<int> mark((size_t)nblk, -1);
  for (int i = 0; i < nrow; i++) {
    DFM_CHECK(rowptr[i + 1] >= rowptr[i], DFM_ERR_ARG, "dfm_sys_create_pattern: rowptr must not decrease");
    for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
      const unsigned j = colind[q];
      DFM_CHECK(j < (unsigned)nblk, DFM_ERR_ARG, "dfm_sys_create_pattern: column out of range");
      DFM_CHECK(j != (unsigned)i, DFM_ERR_ARG, "dfm_sys_create_pattern: the diagonal block is kept apart");
      DFM_CHECK(mark[j] != i, DFM_ERR_ARG, "dfm_sys_create_pattern: a column repeats in a row");
      mark[j] = i;
      ci.push_back(j);
    }
    rp[i + 1] = rowptr[i + 1];
  }
  for (int i = nrow; i < nblk; i++) rp[i + 1] = rp[i];
  dfm_sys h = new dfm_sys_s;
  Sys& s = h->s;
  s.ctx = &ctx->c;
  s.ndim = 0; s.n = nblk; s.len = len; s.bs = len * len; s.npoel = 0; s.nelem = 0;
  s.rowptr0 = std::move(rp);
  s.colind0 = std::move(ci);
  s.elsup_ptr.assign((size_t)nblk + 1, 0);
  s.ncrs = s.rowptr0[nblk];
  std::vector<int> ident(nblk);
  std::iota(ident.begin(), ident.end(), 0);
  int rc = apply_order(s, ident, {}, {}, 0);
  if (rc != DFM_OK) { delete h; return rc; }
  const hipError_t e = sys_alloc_device(s);
  if (e != hipSuccess) {
    set_error(std::string("dfm_sys_create_pattern: ") + hipGetErrorString(e));
    delete h;
    return e == hipErrorOutOfMemory ? DFM_ERR_NOMEM : DFM_ERR_HIP;
  }
  h->owner = ctx;
  ctx->c.refs++;
  *out = h;
  return DFM_OK;
}

static void copy_sweep_host(Sweep& d, const Sweep& a) {
  d.phase_gptr = a.phase_gptr; d.phase_nlev = a.phase_nlev; d.grp = a.grp; d.rowlist = a.rowlist; d.lev = a.lev;
  d.phase_wo = a.phase_wo; d.phase_wi = a.phase_wi;
  d.gmeta = a.gmeta; d.srow = a.srow; d.scnt = a.scnt; d.ecol = a.ecol; d.esrc = a.esrc; d.nent = a.nent;
  d.fuse_bgrp = a.fuse_bgrp;
}

int dfm_sys_clone(dfm_sys src, dfm_sys* out) {
  NEED(src, "dfm_sys_clone");
  NEED(out, "dfm_sys_clone");
  const Sys& a = src->s;
  dfm_sys h = new dfm_sys_s;
  Sys& s = h->s;
  s.ctx = a.ctx;
  s.ndim = a.ndim; s.n = a.n; s.len = a.len; s.bs = a.bs; s.npoel = a.npoel; s.nelem = a.nelem; s.ncrs = a.ncrs;
  s.coords = a.coords; s.lnods = a.lnods; s.rowptr0 = a.rowptr0; s.colind0 = a.colind0;
  s.elsup_ptr = a.elsup_ptr; s.elsup = a.elsup;
  s.order = a.order; s.inv = a.inv; s.natural = a.natural; s.ncolors = a.ncolors;
  s.color_of = a.color_of; s.tile_of = a.tile_of;
  s.rowptr = a.rowptr; s.colind = a.colind; s.diaind = a.diaind; s.lsplit = a.lsplit; s.usplit = a.usplit;
  s.max_rowlen = a.max_rowlen;
  s.fill_level = a.fill_level; s.frowptr = a.frowptr; s.fcolind = a.fcolind; s.fdiaind = a.fdiaind; s.a2f = a.a2f;
  s.perm_crs = a.perm_crs;
  s.eptr = a.eptr; s.eidx = a.eidx; s.eloc = a.eloc;
  s.bc0 = a.bc0;
  copy_sweep_host(s.fwd, a.fwd);
  copy_sweep_host(s.bwd, a.bwd);
  s.sell.sptr = a.sell.sptr; s.sell.col = a.sell.col; s.sell.src = a.sell.src; s.sell.nslab = a.sell.nslab;
  s.sell.visit = a.sell.visit;
  s.fplan.ok = a.fplan.ok; s.fplan.gpass = a.fplan.gpass; s.fplan.pmeta = a.fplan.pmeta; s.fplan.etab = a.fplan.etab;
  s.fplan.utab = a.fplan.utab; s.fplan.ukj = a.fplan.ukj; s.fplan.ustep = a.fplan.ustep;
  auto fail = [&](hipError_t e) {
    set_error(std::string("dfm_sys_clone: ") + hipGetErrorString(e));
    delete h;
    return e == hipErrorOutOfMemory ? DFM_ERR_NOMEM : DFM_ERR_HIP;
  };
  const int rc = upload_structure(s);
  if (rc != DFM_OK) {
    delete h;
    return rc;
  }
  hipError_t e = sys_alloc_device(s);
  if (e != hipSuccess) return fail(e);
  e = hipMemcpyAsync(s.d_bc.p, a.d_bc.p, sizeof(int) * mdof(s), hipMemcpyDeviceToDevice, s.ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s.ctx->stream);
  if (e != hipSuccess) return fail(e);
  h->owner = src->owner;
  if (h->owner) h->owner->c.refs++;
  *out = h;
  return DFM_OK;
}

int dfm_sys_destroy(dfm_sys sys) {
  if (!sys) return DFM_OK;
  if (sys->s.batch) batch_detach(sys->s.batch, &sys->s);
  (void)hipStreamSynchronize(sys->s.ctx->stream);
  dfm_ctx owner = sys->owner;
  delete sys;
  if (owner) ctx_release(owner);
  return DFM_OK;
}

int dfm_sys_info(dfm_sys sys, int* nblk, int* len, long long* ncrs) {
  NEED(sys, "dfm_sys_info");
  if (nblk) *nblk = sys->s.n;
  if (len) *len = sys->s.len;
  if (ncrs) *ncrs = sys->s.ncrs;
  return DFM_OK;
}

int dfm_sys_get_pattern(dfm_sys sys, unsigned* rowptr, unsigned* colind) {
  NEED(sys, "dfm_sys_get_pattern");
  const Sys& s = sys->s;
  if (rowptr) std::memcpy(rowptr, s.rowptr0.data(), sizeof(unsigned) * (s.n + 1));
  if (colind) std::memcpy(colind, s.colind0.data(), sizeof(unsigned) * s.ncrs);
  return DFM_OK;
}

// the last ordering built by dfm_sys_make_ordering, with its colouring (kept so that dfm_sys_set_ordering
// with that same order can use contiguous groups)
struct Pending {
  std::vector<int> order, color, tile;
  int ncolors = 0;
};
static thread_local Pending g_pending;

int dfm_sys_make_ordering(dfm_sys sys, int kind, int tile, int* order_out, int* ncolors_out) {
  NEED(sys, "dfm_sys_make_ordering");
  Sys& s = sys->s;
  Pending p;
  if (kind == DFM_ORDER_NATURAL) {
    p.order.resize(s.n);
    std::iota(p.order.begin(), p.order.end(), 0);
  } else if (kind == DFM_ORDER_MULTICOLOR) {
    make_multicolor(s, p.order, p.color, p.ncolors);
  } else if (kind == DFM_ORDER_BLOCKCOLOR) {
    DFM_CHECK(!s.coords.empty(), DFM_ERR_ARG, "dfm_sys_make_ordering: the block-colour tiles need node coordinates "
                                              "(a system made from a block pattern has none)");
    make_blockcolor(s, tile, p.order, p.color, p.tile, p.ncolors);
  } else if (kind == DFM_ORDER_RCM || kind == DFM_ORDER_RCM2 || kind == DFM_ORDER_AMD) {
    DFM_CHECK(make_reference_ordering(kind, s.n, s.rowptr0.data(), s.colind0.data(), p.order), DFM_ERR_ARG,
              "dfm_sys_make_ordering: the block graph is in several pieces (the reference asserts)");
  } else {
    DFM_CHECK(false, DFM_ERR_ARG, "dfm_sys_make_ordering: unknown kind");
  }
  if (order_out) std::memcpy(order_out, p.order.data(), sizeof(int) * s.n);
  if (ncolors_out) *ncolors_out = p.ncolors;
  g_pending = std::move(p);
  return DFM_OK;
}

int dfm_pattern_ordering(int kind, int n, const unsigned* rowptr, const unsigned* colind, int* order) {
  DFM_CHECK(n >= 0 && (n == 0 || (rowptr && colind && order)), DFM_ERR_ARG, "dfm_pattern_ordering: bad arguments");
  DFM_CHECK(kind == DFM_ORDER_NATURAL || kind == DFM_ORDER_RCM || kind == DFM_ORDER_RCM2 || kind == DFM_ORDER_AMD,
            DFM_ERR_ARG, "dfm_pattern_ordering: kind must be NATURAL, RCM, RCM2 or AMD");
  for (int i = 0; i < n; i++)
    for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++)
      DFM_CHECK(colind[q] < (unsigned)n, DFM_ERR_ARG, "dfm_pattern_ordering: column out of range");
  std::vector<int> ord;
  DFM_CHECK(make_reference_ordering(kind, n, rowptr, colind, ord), DFM_ERR_ARG,
            "dfm_pattern_ordering: the block graph is in several pieces (the reference asserts)");
  if (n) std::memcpy(order, ord.data(), sizeof(int) * n);
  return DFM_OK;
}

int dfm_sys_set_ordering(dfm_sys sys, const int* order) {
  NEED(sys, "dfm_sys_set_ordering");
  Sys& s = sys->s;
  std::vector<int> ord(s.n);
  if (order) {
    std::vector<char> seen(s.n, 0);
    for (int i = 0; i < s.n; i++) {
      DFM_CHECK(order[i] >= 0 && order[i] < s.n && !seen[order[i]], DFM_ERR_ARG,
                "dfm_sys_set_ordering: not a permutation");
      seen[order[i]] = 1;
      ord[i] = order[i];
    }
  } else {
    std::iota(ord.begin(), ord.end(), 0);
  }
  std::vector<int> color, tile;
  int nc = 0;
  if (order && g_pending.order == ord && !g_pending.color.empty()) {
    color = g_pending.color;
    tile = g_pending.tile;
    nc = g_pending.ncolors;
  }
  // Dirichlet flags follow the relabelling
  std::vector<int> bc((size_t)mdof(s));
  DFM_HIP(hipMemcpy(bc.data(), s.d_bc.p, bc.size() * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> bc0(bc.size());
  for (int i0 = 0; i0 < s.n; i0++)
    for (int d = 0; d < s.len; d++) bc0[(size_t)s.order[i0] * s.len + d] = bc[(size_t)i0 * s.len + d];
  const int rc = apply_order(s, ord, color, tile, nc);
  if (rc != DFM_OK) return rc;
  for (int i0 = 0; i0 < s.n; i0++)
    for (int d = 0; d < s.len; d++) bc[(size_t)i0 * s.len + d] = bc0[(size_t)s.order[i0] * s.len + d];
  DFM_HIP(hipMemcpy(s.d_bc.p, bc.data(), bc.size() * sizeof(int), hipMemcpyHostToDevice));
  return DFM_OK;
}

int dfm_sys_schedule_info(dfm_sys sys, int* nphase_fwd, int* nphase_bwd, int* max_group_levels,
                          int* pcg_sweep_launches) {
  NEED(sys, "dfm_sys_schedule_info");
  const Sys& s = sys->s;
  if (nphase_fwd) *nphase_fwd = s.fwd.nphase();
  if (nphase_bwd) *nphase_bwd = s.bwd.nphase();
  if (pcg_sweep_launches) *pcg_sweep_launches = s.fwd.nphase() + s.bwd.nphase() - (s.fwd.fuse_bgrp.empty() ? 0 : 1);
  if (max_group_levels) {
    int mx = 0;
    for (int v : s.fwd.phase_nlev) mx = std::max(mx, v);
    for (int v : s.bwd.phase_nlev) mx = std::max(mx, v);
    *max_group_levels = mx;
  }
  return DFM_OK;
}

int dfm_sys_set_bc(dfm_sys sys, const int* flags) {
  NEED(sys, "dfm_sys_set_bc");
  NEED(flags, "dfm_sys_set_bc");
  Sys& s = sys->s;
  s.bc0.assign((size_t)mdof(s), 0);
  for (long long k = 0; k < mdof(s); k++) s.bc0[(size_t)k] = flags[k] ? 1 : 0;
  s.saved = SAVE_NONE;  // the boundary coupling depends on the flags
  return upload_bc(s);
}

int dfm_sys_set_field(dfm_sys sys, const double* u, const double* v) {
  NEED(sys, "dfm_sys_set_field");
  Sys& s = sys->s;
  const size_t bytes = sizeof(double) * mdof(s);
  hipStream_t st = s.ctx->stream;
  // on the library's stream (non-blocking: a null-stream copy would not wait for an assembly still reading u, v)
  if (u) DFM_HIP(hipMemcpyAsync(s.d_u.p, u, bytes, hipMemcpyHostToDevice, st));
  else DFM_HIP(launch_fill(s.d_u.p, 0.0, mdof(s), st));
  if (v) DFM_HIP(hipMemcpyAsync(s.d_v.p, v, bytes, hipMemcpyHostToDevice, st));
  else DFM_HIP(launch_fill(s.d_v.p, 0.0, mdof(s), st));
  DFM_HIP(hipStreamSynchronize(st));
  return DFM_OK;
}

int dfm_sys_get_field(dfm_sys sys, double* u) {
  NEED(sys, "dfm_sys_get_field");
  NEED(u, "dfm_sys_get_field");
  Sys& s = sys->s;
  hipStream_t st = s.ctx->stream;
  DFM_HIP(hipMemcpyAsync(u, s.d_u.p, sizeof(double) * mdof(s), hipMemcpyDeviceToHost, st));
  DFM_HIP(hipStreamSynchronize(st));
  return DFM_OK;
}

int dfm_sys_update_field(dfm_sys sys) {
  NEED(sys, "dfm_sys_update_field");
  Sys& s = sys->s;
  DFM_HIP(launch_update_field(s.n, s.len, s.d_order.p, s.d_upd.p, s.d_u.p, s.ctx->stream));
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  return DFM_OK;
}

// the ACCELERATION field, zero until set
static int ensure_acc(Sys& s) {
  if (s.d_a.p) return DFM_OK;
  DFM_HIP(s.d_a.alloc((size_t)mdof(s)));
  DFM_HIP(launch_fill(s.d_a.p, 0.0, mdof(s), s.ctx->stream));
  return DFM_OK;
}

// the previous velocity of the backward-Euler loop, zero until set
static int ensure_v0(Sys& s) {
  if (s.d_v0.p) return DFM_OK;
  DFM_HIP(s.d_v0.alloc((size_t)mdof(s)));
  DFM_HIP(launch_fill(s.d_v0.p, 0.0, mdof(s), s.ctx->stream));
  return DFM_OK;
}

int dfm_sys_set_velocity_prev(dfm_sys sys, const double* v0) {
  NEED(sys, "dfm_sys_set_velocity_prev");
  Sys& s = sys->s;
  int rc = ensure_v0(s);
  if (rc != DFM_OK) return rc;
  if (v0) DFM_HIP(hipMemcpyAsync(s.d_v0.p, v0, sizeof(double) * mdof(s), hipMemcpyHostToDevice, s.ctx->stream));
  else DFM_HIP(launch_fill(s.d_v0.p, 0.0, mdof(s), s.ctx->stream));
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  return DFM_OK;
}

int dfm_sys_update_field_beuler(dfm_sys sys, double dt, int is_initial) {
  NEED(sys, "dfm_sys_update_field_beuler");
  Sys& s = sys->s;
  DFM_HIP(launch_update_field_beuler(s.n, s.len, s.d_order.p, s.d_upd.p, dt, is_initial ? 1 : 0, s.d_u.p, s.d_v.p,
                                     s.ctx->stream));
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  return DFM_OK;
}

int dfm_sys_update_field_newmark1(dfm_sys sys, double gamma, double dt, int is_initial) {
  NEED(sys, "dfm_sys_update_field_newmark1");
  Sys& s = sys->s;
  const double gdt = gamma * dt;  // the reference's argument gamma*dt, formed before the products
  DFM_HIP(launch_update_field_newmark1(s.n, s.len, s.d_order.p, s.d_upd.p, gdt, dt, is_initial ? 1 : 0, s.d_u.p,
                                       s.d_v.p, s.ctx->stream));
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  return DFM_OK;
}

int dfm_sys_set_acceleration(dfm_sys sys, const double* a) {
  NEED(sys, "dfm_sys_set_acceleration");
  Sys& s = sys->s;
  int rc = ensure_acc(s);
  if (rc != DFM_OK) return rc;
  // on the library's stream, after the zero fill of a first use (the stream does not wait for the null stream)
  if (a) DFM_HIP(hipMemcpyAsync(s.d_a.p, a, sizeof(double) * mdof(s), hipMemcpyHostToDevice, s.ctx->stream));
  else DFM_HIP(launch_fill(s.d_a.p, 0.0, mdof(s), s.ctx->stream));
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  return DFM_OK;
}

int dfm_sys_get_fields(dfm_sys sys, double* u, double* v, double* a) {
  NEED(sys, "dfm_sys_get_fields");
  Sys& s = sys->s;
  const size_t bytes = sizeof(double) * mdof(s);
  hipStream_t st = s.ctx->stream;
  if (u) DFM_HIP(hipMemcpyAsync(u, s.d_u.p, bytes, hipMemcpyDeviceToHost, st));
  if (v) DFM_HIP(hipMemcpyAsync(v, s.d_v.p, bytes, hipMemcpyDeviceToHost, st));
  if (a) {
    if (s.d_a.p) DFM_HIP(hipMemcpyAsync(a, s.d_a.p, bytes, hipMemcpyDeviceToHost, st));
    else std::memset(a, 0, bytes);
  }
  DFM_HIP(hipStreamSynchronize(st));
  return DFM_OK;
}

int dfm_sys_update_field_newmark(dfm_sys sys, double gamma, double beta, double dt, int is_initial) {
  NEED(sys, "dfm_sys_update_field_newmark");
  Sys& s = sys->s;
  int rc = ensure_acc(s);
  if (rc != DFM_OK) return rc;
  DFM_HIP(launch_update_field_newmark(s.n, s.len, s.d_order.p, s.d_upd.p, gamma, beta, dt, is_initial ? 1 : 0,
                                      s.d_u.p, s.d_v.p, s.d_a.p, s.ctx->stream));
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  return DFM_OK;
}

int dfm_sys_set_params(dfm_sys sys, const double* params, int npar, int pstride) {
  NEED(sys, "dfm_sys_set_params");
  NEED(params, "dfm_sys_set_params");
  Sys& s = sys->s;
  DFM_CHECK(npar > 0 && (pstride == 0 || pstride >= npar), DFM_ERR_ARG, "dfm_sys_set_params: bad npar/pstride");
  const size_t npv = pstride ? (size_t)s.nelem * pstride : (size_t)npar;
  if (s.d_par.n != npv) DFM_HIP(s.d_par.alloc(npv));
  DFM_HIP(hipMemcpy(s.d_par.p, params, npv * sizeof(double), hipMemcpyHostToDevice));
  s.npar = npar;
  s.pstride = pstride;
  return DFM_OK;
}

// consistent P1 surface loads t*A/3 per node (a reference extension), kept per node in face order so that
// the assembly adds them after the element contributions in a fixed order
int dfm_sys_set_loads(dfm_sys sys, int nface, const int* faces, const double* trac) {
  NEED(sys, "dfm_sys_set_loads");
  Sys& s = sys->s;
  if (nface <= 0) {
    s.has_loads = false;
    return DFM_OK;
  }
  DFM_CHECK(s.ndim == 3 && s.len == 3 && faces && trac, DFM_ERR_ARG,
            "dfm_sys_set_loads: surface loads need a 3D vector field, faces and tractions");
  std::vector<int> lptr(s.n + 1, 0);
  for (long long f = 0; f < nface; f++)
    for (int k = 0; k < 3; k++) {
      const int nd = faces[f * 3 + k];
      DFM_CHECK(nd >= 0 && nd < s.n, DFM_ERR_ARG, "dfm_sys_set_loads: face node out of range");
      lptr[nd + 1]++;
    }
  for (int i = 0; i < s.n; i++) lptr[i + 1] += lptr[i];
  std::vector<double> loads((size_t)lptr[s.n] * 3, 0.0);
  std::vector<int> fill(lptr.begin(), lptr.end() - 1);
  for (long long f = 0; f < nface; f++) {
    const double* p0 = &s.coords[(size_t)faces[f * 3 + 0] * 3];
    const double* p1 = &s.coords[(size_t)faces[f * 3 + 1] * 3];
    const double* p2 = &s.coords[(size_t)faces[f * 3 + 2] * 3];
    const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double area = 0.5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int k = 0; k < 3; k++) {
      const int slot = fill[faces[f * 3 + k]]++;
      for (int d = 0; d < 3; d++) loads[(size_t)slot * 3 + d] = trac[f * 3 + d] * area / 3.0;
    }
  }
  hipStream_t st = s.ctx->stream;
  DFM_HIP(upload(s.d_lptr, lptr, st));
  DFM_HIP(upload(s.d_loads, loads, st));
  s.has_loads = true;
  return DFM_OK;
}

int dfm_sys_set_load_scale(dfm_sys sys, double scale) {
  NEED(sys, "dfm_sys_set_load_scale");
  sys->s.load_scale = scale;
  return DFM_OK;
}

static int check_kind(const Sys& s, int kind) {
  DFM_CHECK(s.nelem > 0, DFM_ERR_STATE, "dfm_sys_assemble: a system made from a block pattern has no elements "
                                        "(its values come from dfm_sys_set_matrix)");
  static const int need_npar[12] = {2, 5, 6, 5, 6, 3, 5, 6, 5, 6, 5, 5};
  static const int need_len[12] = {1, 2, 3, 2, 3, 1, 2, 3, 2, 3, 2, 2};
  DFM_CHECK(kind >= 0 && kind < 12, DFM_ERR_ARG, "dfm_sys_assemble: unknown element kind");
  DFM_CHECK(s.d_par.p != nullptr, DFM_ERR_STATE, "dfm_sys_assemble: no element parameters set");
  DFM_CHECK(s.npar >= need_npar[kind], DFM_ERR_ARG, "dfm_sys_assemble: too few parameters for the element kind");
  DFM_CHECK(need_len[kind] == s.len, DFM_ERR_ARG, "dfm_sys_assemble: element kind does not match dofs per node");
  const bool k3 = kind == DFM_SOLID3D || kind == DFM_STVK3D || kind == DFM_SOLID3D_NEWMARK || kind == DFM_STVK3D_NEWMARK;
  DFM_CHECK(k3 || kind == DFM_POISSON || s.ndim == 2, DFM_ERR_ARG, "dfm_sys_assemble: element kind needs a 2D mesh");
  DFM_CHECK(!k3 || s.ndim == 3, DFM_ERR_ARG, "dfm_sys_assemble: element kind needs a 3D mesh");
  return DFM_OK;
}

// SELL copy of the matrix for the Krylov SpMV; the system is assembled and needs a new factor
static int finish_matrix(Sys& s, hipStream_t st) {
  const long long ne = s.sell.nslab * 64;
  if (s.sell.d_val.n < (size_t)(ne * s.bs)) DFM_HIP(s.sell.d_val.alloc((size_t)std::max<long long>(ne, 1) * s.bs));
  DFM_HIP(launch_slab_fill(s.len, ne, s.sell.d_src.p, s.d_crs.p, s.sell.d_val.p, st));
  s.assembled = true;
  s.factored = s.factored32 = false;
  s.prepared = -1;
  return DFM_OK;
}

}  // extern "C"

// The element store of an assembly of s (its lists and its context's buffer, grown as needed); false where the
// lists are not built or the buffer cannot be had (the per-row gather runs then). The vector kinds (one dof per
// dimension: linear elasticity, StVK, their Newmark and backward-Euler forms) with rows of at most 63 off-diagonal
// blocks take the row-group kernel, which needs neither.
static bool elem_store(Sys& s, int kind, bool save, ElemStore& es) {
  es = ElemStore{};
  const bool vector_kind = kind == DFM_SOLID2D || kind == DFM_SOLID3D || kind == DFM_STVK2D || kind == DFM_STVK3D ||
                           kind == DFM_SOLID2D_NEWMARK || kind == DFM_SOLID3D_NEWMARK || kind == DFM_SOLID2D_BEULER ||
                           kind == DFM_STVK3D_NEWMARK || kind == DFM_STVK2D_NEWMARK || kind == DFM_STVK2D_BEULER;
  if (!save && vector_kind && s.len == s.ndim && std::getenv("DFM_ASM_ELEMSTORE") == nullptr && s.nelem > 0 &&
      s.max_rowlen < 64 && s.d_permq.p && s.d_rowptr0.p && s.d_colind0.p && s.npoel <= 4) {
    if (!s.d_elnod.p) {
      std::vector<int> en(s.elsup.size() * 4, -1);
      for (size_t k = 0; k < s.elsup.size(); k++)
        for (int c = 0; c < s.npoel; c++) en[k * 4 + c] = s.lnods[(size_t)s.elsup[k] * s.npoel + c];
      if (upload(s.d_elnod, en, s.ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        s.d_elnod.release();
        return false;
      }
    }
    int lpr = 8;
    while ((unsigned)lpr < s.max_rowlen + 1) lpr *= 2;
    es.nelem = s.nelem;
    es.ncrs = s.ncrs;
    es.lpr = lpr;
    return true;
  }
  if (s.eptr.empty() || !s.d_eptr.p || !s.d_eloc.p) return false;
  const size_t need = (size_t)s.nelem * s.npoel * ((size_t)s.npoel * s.bs + s.len);
  DBuf<double>& b = s.ctx->ebuf;
  if (b.n < need) {
    // the context's stream is the only user of the buffer: wait for it before the old one goes back to the cache
    if (hipStreamSynchronize(s.ctx->stream) != hipSuccess) return false;
    {
      CacheRelease guard;
      b.release();
    }
    if (b.alloc(need) != hipSuccess) {
      (void)hipGetLastError();
      b.release();
      return false;
    }
  }
  es.nelem = s.nelem;
  es.ncrs = s.ncrs;
  es.eptr = s.d_eptr.p;
  es.eidx = s.d_eidx.p;
  es.eloc = s.d_eloc.p;
  es.ebuf = b.p;
  return true;
}

extern "C" {

int dfm_sys_assemble_resident(dfm_sys sys, int kind, const double* glob, double* res_norm) {
  NEED(sys, "dfm_sys_assemble_resident");
  Sys& s = sys->s;
  int rc = check_kind(s, kind);
  if (rc != DFM_OK) return rc;
  const bool newmark = kind == DFM_SOLID2D_NEWMARK || kind == DFM_SOLID3D_NEWMARK || kind == DFM_STVK3D_NEWMARK ||
                       kind == DFM_STVK2D_NEWMARK;
  if (newmark) {
    DFM_CHECK(glob != nullptr, DFM_ERR_ARG, "dfm_sys_assemble: Newmark kinds need glob = (dt, gamma, beta, is_initial)");
    rc = ensure_acc(s);
    if (rc != DFM_OK) return rc;
  }
  if (kind == DFM_SOLID2D_BEULER || kind == DFM_STVK2D_BEULER) {
    DFM_CHECK(glob != nullptr, DFM_ERR_ARG, "dfm_sys_assemble: the backward-Euler kind needs glob = (dt, is_initial)");
    rc = ensure_v0(s);
    if (rc != DFM_OK) return rc;
  }
  hipStream_t st = s.ctx->stream;
  const long long m = mdof(s);
  ElemStore es{};
  const bool store = elem_store(s, kind, false, es);
  // InitializeMarge (src/ls/linearsystem.cpp:74-88): matrix, residual and update set to zero (the element store
  // writes every entry of the matrix and the residual itself)
  if (!store) {
    DFM_HIP(launch_fill(s.d_crs.p, 0.0, (long long)s.ncrs * s.bs, st));
    DFM_HIP(launch_fill(s.d_dia.p, 0.0, (long long)s.n * s.bs, st));
    DFM_HIP(launch_fill(s.d_res.p, 0.0, m, st));
  }
  DFM_HIP(launch_fill(s.d_upd.p, 0.0, m, st));
  AsmArgs a{};
  a.kind = kind; a.ndim = s.ndim; a.npoel = s.npoel; a.len = s.len; a.n = s.n;
  a.coords = s.d_coords.p; a.lnods = s.d_lnods.p; a.elsup_ptr = s.d_elsup_ptr.p; a.elsup = s.d_elsup.p;
  a.order = s.d_order.p; a.inv = s.d_inv.p;
  a.params = s.d_par.p; a.npar = s.npar; a.pstride = s.pstride;
  // only the kind's own scalars are read (kind_glob_count): a caller's array may hold no more
  const int ng = glob ? kind_glob_count(kind) : 0;
  a.glob0 = ng > 0 ? glob[0] : 0.0; a.glob1 = ng > 1 ? glob[1] : 0.0;
  a.glob2 = ng > 2 ? glob[2] : 0.0; a.glob3 = ng > 3 ? glob[3] : 0.0;
  a.u = s.d_u.p; a.v = s.d_v.p; a.acc = (kind == DFM_SOLID2D_BEULER || kind == DFM_STVK2D_BEULER) ? s.d_v0.p : s.d_a.p;
  a.load_ptr = s.has_loads ? s.d_lptr.p : nullptr;
  a.loads = s.has_loads ? s.d_loads.p : nullptr;
  a.load_scale = s.load_scale;
  a.rowptr = s.d_rowptr.p; a.colind = s.d_colind.p; a.crs = s.d_crs.p; a.dia = s.d_dia.p; a.res = s.d_res.p;
  a.rowptr0 = s.d_rowptr0.p; a.colind0 = s.d_colind0.p; a.permq = s.d_permq.p; a.elnod = s.d_elnod.p;
  DFM_HIP(launch_assemble(a, store ? &es : nullptr, st));
  if (!s.last_asm) s.last_asm = new AsmArgs;
  *s.last_asm = a;
  // FinalizeMarge (:92-128): Dirichlet rows/columns, residual zeroed at fixed dofs, ||r|| returned
  DFM_HIP(launch_set_bc(s.n, s.len, s.d_rowptr.p, s.d_colind.p, s.d_crs.p, s.d_dia.p, s.d_res.p, s.d_bc.p, st));
  DFM_HIP(launch_copy(s.d_res.p, s.d_rhs.p, m, st));
  rc = finish_matrix(s, st);
  if (rc != DFM_OK) return rc;
  s.saved = SAVE_NONE;
  if (res_norm) {
    DFM_HIP(launch_dot(s.d_res.p, s.d_res.p, m, s.d_partial.p, s.d_counter.p, s.d_scal.p + S_TMP, st));
    DFM_HIP(hipMemcpyAsync(s.h_pinned_d, s.d_scal.p + S_TMP, sizeof(double), hipMemcpyDeviceToHost, st));
    DFM_HIP(hipStreamSynchronize(st));
    *res_norm = std::sqrt(s.h_pinned_d[0]);
  }
  return DFM_OK;
}

int dfm_sys_assemble(dfm_sys sys, int kind, const double* params, int npar, int pstride, const double* glob,
                     int nface, const int* faces, const double* trac, double* res_norm) {
  NEED(sys, "dfm_sys_assemble");
  NEED(params, "dfm_sys_assemble");
  int rc = dfm_sys_set_params(sys, params, npar, pstride);
  if (rc != DFM_OK) return rc;
  rc = dfm_sys_set_loads(sys, nface, faces, trac);
  if (rc != DFM_OK) return rc;
  rc = dfm_sys_assemble_resident(sys, kind, glob, res_norm);
  if (rc != DFM_OK) return rc;
  DFM_HIP(hipStreamSynchronize(sys->s.ctx->stream));
  return DFM_OK;
}

// The matrix values from the host (the layout of dfm_sys_get_matrix: the caller's blocks in pattern order, the
// diagonal blocks by node), as the element routines leave them in a CMatDia_BlkCrs through
// ILinearSystem_Eqn::GetMatrix (include/delfem/linearsystem_interface_eqnsys.h:48-68); the system counts as
// assembled (its factor is stale).
int dfm_sys_set_matrix(dfm_sys sys, const double* crs, const double* dia) {
  NEED(sys, "dfm_sys_set_matrix");
  NEED(dia, "dfm_sys_set_matrix");
  Sys& s = sys->s;
  DFM_CHECK(crs || s.ncrs == 0, DFM_ERR_ARG, "dfm_sys_set_matrix: null crs");
  hipStream_t st = s.ctx->stream;
  std::vector<double> c((size_t)s.ncrs * s.bs), d((size_t)s.n * s.bs);
  for (long long q = 0; q < s.ncrs; q++)
    std::memcpy(&c[(size_t)s.perm_crs[q] * s.bs], &crs[(size_t)q * s.bs], sizeof(double) * s.bs);
  for (int i0 = 0; i0 < s.n; i0++)
    std::memcpy(&d[(size_t)i0 * s.bs], &dia[(size_t)s.order[i0] * s.bs], sizeof(double) * s.bs);
  DFM_HIP(hipStreamSynchronize(st));
  if (!c.empty()) DFM_HIP(hipMemcpyAsync(s.d_crs.p, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice, st));
  DFM_HIP(hipMemcpyAsync(s.d_dia.p, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice, st));
  DFM_HIP(hipStreamSynchronize(st));  // the host arrays go out of scope
  const int rc = finish_matrix(s, st);
  if (rc != DFM_OK) return rc;
  s.saved = SAVE_NONE;
  DFM_HIP(hipStreamSynchronize(st));
  return DFM_OK;
}

}  // extern "C"

// FinalizeMarge of a system whose values came from the host: Dirichlet rows and columns by the flags of
// dfm_sys_set_bc, the residual zeroed at fixed dofs and kept as the right-hand side, the SELL values refreshed.
int dfm::sys_apply_bc(Sys& s) {
  hipStream_t st = s.ctx->stream;
  DFM_HIP(launch_set_bc(s.n, s.len, s.d_rowptr.p, s.d_colind.p, s.d_crs.p, s.d_dia.p, s.d_res.p, s.d_bc.p, st));
  DFM_HIP(launch_copy(s.d_res.p, s.d_rhs.p, mdof(s), st));
  return finish_matrix(s, st);
}

extern "C" {

int dfm_sys_get_matrix(dfm_sys sys, double* crs, double* dia) {
  NEED(sys, "dfm_sys_get_matrix");
  Sys& s = sys->s;
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  std::vector<double> c((size_t)s.ncrs * s.bs), d((size_t)s.n * s.bs);
  DFM_HIP(hipMemcpy(c.data(), s.d_crs.p, c.size() * sizeof(double), hipMemcpyDeviceToHost));
  DFM_HIP(hipMemcpy(d.data(), s.d_dia.p, d.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (crs)
    for (long long q = 0; q < s.ncrs; q++)
      std::memcpy(&crs[(size_t)q * s.bs], &c[(size_t)s.perm_crs[q] * s.bs], sizeof(double) * s.bs);
  if (dia)
    for (int i0 = 0; i0 < s.n; i0++)
      std::memcpy(&dia[(size_t)s.order[i0] * s.bs], &d[(size_t)i0 * s.bs], sizeof(double) * s.bs);
  return DFM_OK;
}

// K_b as scalar terms per device dof: for a free dof, the fixed columns of its diagonal block, then those of
// each block of its row in pattern order. This is the boundary matrix of the reference after its Dirichlet
// treatment (CMat_BlkCrs::SetPatternBoundary, src/matvec/mat_blkcrs.cpp:1177-1265, keeps the diagonal block
// first and the blocks with a fixed column; SetBoundaryConditionInverse_Colum, :876-900, zeroes the free
// columns; SetBoundaryCondition_Row, :910-953, the fixed rows), less the entries that are zero by that treatment.
//
// full_rows: the lumped-mass saved systems' boundary matrix instead (CMat_BlkCrs::SetPatternDia,
// src/matvec/mat_blkcrs.cpp:1267-1319: the whole row of the matrix pattern, the diagonal block placed before the
// first column below the row, last in a row with none; only its fixed rows are zeroed, :910-953): for a free dof,
// every column of those blocks in that order.
static int build_boundary_terms(Sys& s, bool full_rows = false) {
  const int B = s.len, bs = s.bs;
  const long long m = mdof(s), ndia = (long long)s.n * bs;
  std::vector<int> fl = s.bc0;
  if (fl.empty()) fl.assign((size_t)m, 0);
  std::vector<unsigned> bptr((size_t)m + 1, 0);
  std::vector<unsigned> bcol;
  std::vector<long long> bsrc;
  for (int i0 = 0; i0 < s.n; i0++) {
    const int i1 = s.order[i0];
    for (int d = 0; d < B; d++) {
      const long long k = (long long)i0 * B + d;
      if (!fl[(size_t)i1 * B + d] && full_rows) {
        bool dia_done = false;
        auto dia_block = [&]() {
          for (int jd = 0; jd < B; jd++) {
            bcol.push_back((unsigned)((long long)i1 * B + jd));
            bsrc.push_back((long long)i0 * bs + d * B + jd);
          }
          dia_done = true;
        };
        for (unsigned q = s.rowptr0[i1]; q < s.rowptr0[i1 + 1]; q++) {
          const unsigned j1 = s.colind0[q];
          if ((unsigned)i1 > j1 && !dia_done) dia_block();
          for (int jd = 0; jd < B; jd++) {
            bcol.push_back((unsigned)((long long)j1 * B + jd));
            bsrc.push_back(ndia + s.perm_crs[q] * bs + d * B + jd);
          }
        }
        if (!dia_done) dia_block();
      } else if (!fl[(size_t)i1 * B + d]) {
        for (int jd = 0; jd < B; jd++)
          if (fl[(size_t)i1 * B + jd]) {
            bcol.push_back((unsigned)((long long)i1 * B + jd));
            bsrc.push_back((long long)i0 * bs + d * B + jd);
          }
        for (unsigned q = s.rowptr0[i1]; q < s.rowptr0[i1 + 1]; q++) {
          const unsigned j1 = s.colind0[q];
          for (int jd = 0; jd < B; jd++)
            if (fl[(size_t)j1 * B + jd]) {
              bcol.push_back((unsigned)((long long)j1 * B + jd));
              bsrc.push_back(ndia + s.perm_crs[q] * bs + d * B + jd);
            }
        }
      }
      DFM_CHECK(bcol.size() < 0xffffffffull, DFM_ERR_ARG, "dfm_sys_assemble_save: too many boundary terms");
      bptr[(size_t)k + 1] = (unsigned)bcol.size();
    }
  }
  hipStream_t st = s.ctx->stream;
  s.nbterm = (long long)bcol.size();
  DFM_HIP(upload(s.d_bptr, bptr, st));
  DFM_HIP(upload(s.d_bcol, bcol, st));
  DFM_HIP(upload(s.d_bsrc, bsrc, st));
  if (s.d_bval.n < (size_t)std::max<long long>(s.nbterm, 1)) DFM_HIP(s.d_bval.alloc((size_t)std::max<long long>(s.nbterm, 1)));
  return DFM_OK;
}

int dfm_sys_assemble_save(dfm_sys sys, int kind, const double* params, int npar, int pstride) {
  NEED(sys, "dfm_sys_assemble_save");
  NEED(params, "dfm_sys_assemble_save");
  Sys& s = sys->s;
  DFM_CHECK(kind == DFM_POISSON || kind == DFM_SOLID2D || kind == DFM_SOLID3D, DFM_ERR_ARG,
            "dfm_sys_assemble_save: the reference saves the stiffness of Poisson, LinearSolid2D and LinearSolid3D only");
  DFM_CHECK(kind != DFM_POISSON || s.ndim == 2, DFM_ERR_ARG,
            "dfm_sys_assemble_save: the reference's saved-stiffness Poisson loop takes triangles only");
  int rc = dfm_sys_set_params(sys, params, npar, pstride);
  if (rc != DFM_OK) return rc;
  rc = check_kind(s, kind);
  if (rc != DFM_OK) return rc;
  hipStream_t st = s.ctx->stream;
  const long long m = mdof(s);
  if (s.d_force.n != (size_t)m) DFM_HIP(s.d_force.alloc((size_t)m));
  // InitializeMarge of CLinearSystem_Save (src/femls/linearsystem_fieldsave.cpp:169-187): matrix, residual,
  // update, force and boundary matrix set to zero
  DFM_HIP(launch_fill(s.d_crs.p, 0.0, (long long)s.ncrs * s.bs, st));
  DFM_HIP(launch_fill(s.d_dia.p, 0.0, (long long)s.n * s.bs, st));
  DFM_HIP(launch_fill(s.d_res.p, 0.0, m, st));
  DFM_HIP(launch_fill(s.d_upd.p, 0.0, m, st));
  DFM_HIP(launch_fill(s.d_force.p, 0.0, m, st));
  AsmArgs a{};
  a.kind = kind; a.ndim = s.ndim; a.npoel = s.npoel; a.len = s.len; a.n = s.n;
  a.coords = s.d_coords.p; a.lnods = s.d_lnods.p; a.elsup_ptr = s.d_elsup_ptr.p; a.elsup = s.d_elsup.p;
  a.order = s.d_order.p; a.inv = s.d_inv.p;
  a.params = s.d_par.p; a.npar = s.npar; a.pstride = s.pstride;
  a.rowptr = s.d_rowptr.p; a.colind = s.d_colind.p; a.crs = s.d_crs.p; a.dia = s.d_dia.p; a.res = s.d_force.p;
  a.save = true;
  ElemStore es{};
  const bool store = elem_store(s, kind, true, es);
  DFM_HIP(launch_assemble(a, store ? &es : nullptr, st));
  rc = build_boundary_terms(s);
  if (rc != DFM_OK) return rc;
  DFM_HIP(launch_gather_terms(s.nbterm, s.d_bsrc.p, (long long)s.n * s.bs, s.d_dia.p, s.d_crs.p, s.d_bval.p, st));
  // FinalizeMarge (:189-204): the matrix's Dirichlet rows and columns (the residual is still zero)
  DFM_HIP(launch_set_bc(s.n, s.len, s.d_rowptr.p, s.d_colind.p, s.d_crs.p, s.d_dia.p, s.d_res.p, s.d_bc.p, st));
  DFM_HIP(launch_copy(s.d_res.p, s.d_rhs.p, m, st));
  rc = finish_matrix(s, st);
  if (rc != DFM_OK) return rc;
  s.saved = SAVE_PLAIN;
  DFM_HIP(hipStreamSynchronize(st));
  return DFM_OK;
}

int dfm_sys_assemble_save_dia(dfm_sys sys, int kind, const double* params, int npar, int pstride,
                              const double* glob) {
  NEED(sys, "dfm_sys_assemble_save_dia");
  NEED(params, "dfm_sys_assemble_save_dia");
  NEED(glob, "dfm_sys_assemble_save_dia");
  Sys& s = sys->s;
  DFM_CHECK((kind == DFM_DIFFUSION2D || kind == DFM_SOLID2D) && s.ndim == 2, DFM_ERR_ARG,
            "dfm_sys_assemble_save_dia: the reference's lumped-mass saved systems take DFM_DIFFUSION2D "
            "(CLinearSystem_SaveDiaM_Newmark) or DFM_SOLID2D (CLinearSystem_SaveDiaM_NewmarkBeta) on triangles");
  int rc = dfm_sys_set_params(sys, params, npar, pstride);
  if (rc != DFM_OK) return rc;
  rc = check_kind(s, kind);
  if (rc != DFM_OK) return rc;
  hipStream_t st = s.ctx->stream;
  const long long m = mdof(s);
  if (s.d_force.n != (size_t)m) DFM_HIP(s.d_force.alloc((size_t)m));
  // InitializeMarge (src/femls/linearsystem_fieldsave.cpp:398-420, :662-686): matrix, residual, update, force and
  // boundary matrix to zero
  auto zero = [&]() -> hipError_t {
    hipError_t e;
    if ((e = launch_fill(s.d_crs.p, 0.0, (long long)s.ncrs * s.bs, st)) != hipSuccess) return e;
    if ((e = launch_fill(s.d_dia.p, 0.0, (long long)s.n * s.bs, st)) != hipSuccess) return e;
    if ((e = launch_fill(s.d_res.p, 0.0, m, st)) != hipSuccess) return e;
    return launch_fill(s.d_upd.p, 0.0, m, st);
  };
  DFM_HIP(zero());
  DFM_HIP(launch_fill(s.d_force.p, 0.0, m, st));
  AsmArgs a{};
  a.ndim = s.ndim; a.npoel = s.npoel; a.len = s.len; a.n = s.n;
  a.coords = s.d_coords.p; a.lnods = s.d_lnods.p; a.elsup_ptr = s.d_elsup_ptr.p; a.elsup = s.d_elsup.p;
  a.order = s.d_order.p; a.inv = s.d_inv.p;
  a.npar = s.npar; a.pstride = s.pstride;
  a.rowptr = s.d_rowptr.p; a.colind = s.d_colind.p; a.crs = s.d_crs.p; a.dia = s.d_dia.p;
  // 1. the boundary matrix's values: the stiffness alone, K (diffusion: alpha area grad grad, the saved Poisson
  // element on (alpha, source); elasticity: the saved static element), in the same element order as the matrix;
  // its force goes to the residual handle, which is cleared again
  a.kind = kind == DFM_DIFFUSION2D ? DFM_POISSON : DFM_SOLID2D;
  a.params = kind == DFM_DIFFUSION2D ? s.d_par.p + 1 : s.d_par.p;
  a.res = s.d_res.p;
  a.save = true;
  ElemStore es{};
  const bool store = elem_store(s, a.kind, true, es);
  DFM_HIP(launch_assemble(a, store ? &es : nullptr, st));
  rc = build_boundary_terms(s, true);
  if (rc != DFM_OK) return rc;
  DFM_HIP(launch_gather_terms(s.nbterm, s.d_bsrc.p, (long long)s.n * s.bs, s.d_dia.p, s.d_crs.p, s.d_bval.p, st));
  // 2. the system matrix and the force: the lumped-mass element loop
  DFM_HIP(zero());
  a.kind = kind == DFM_DIFFUSION2D ? KIND_DIFFUSION2D_SAVEDIA : KIND_SOLID2D_SAVEDIA;
  a.params = s.d_par.p;
  a.glob0 = glob[0];
  a.glob1 = glob[1];
  a.glob2 = kind == DFM_SOLID2D ? glob[2] : 0.0;
  a.res = s.d_force.p;
  a.save = false;
  DFM_HIP(launch_assemble(a, nullptr, st));
  // FinalizeMarge: the matrix's Dirichlet rows and columns (the residual is still zero)
  DFM_HIP(launch_set_bc(s.n, s.len, s.d_rowptr.p, s.d_colind.p, s.d_crs.p, s.d_dia.p, s.d_res.p, s.d_bc.p, st));
  DFM_HIP(launch_copy(s.d_res.p, s.d_rhs.p, m, st));
  rc = finish_matrix(s, st);
  if (rc != DFM_OK) return rc;
  s.saved = kind == DFM_DIFFUSION2D ? SAVE_DIA_NEWMARK : SAVE_DIA_NEWMARKBETA;
  s.save_dt = glob[0];
  s.save_gamma = glob[1];
  s.save_beta = kind == DFM_SOLID2D ? glob[2] : 0.0;
  DFM_HIP(hipStreamSynchronize(st));
  return DFM_OK;
}

int dfm_sys_make_residual(dfm_sys sys, double* res_norm) {
  NEED(sys, "dfm_sys_make_residual");
  Sys& s = sys->s;
  DFM_CHECK(s.saved, DFM_ERR_STATE,
            "dfm_sys_make_residual: assemble with dfm_sys_assemble_save or dfm_sys_assemble_save_dia first");
  hipStream_t st = s.ctx->stream;
  const long long m = mdof(s);
  double c0 = 0.0, c1 = 0.0;
  if (s.saved == SAVE_DIA_NEWMARK) c0 = s.save_dt * (1 - s.save_gamma);
  if (s.saved == SAVE_DIA_NEWMARKBETA) {
    c0 = s.save_dt;
    c1 = s.save_dt * s.save_dt * (0.5 - s.save_beta);
    const int rc = ensure_acc(s);
    if (rc != DFM_OK) return rc;
  }
  DFM_HIP(launch_make_residual(s.n, s.len, s.saved, c0, c1, s.d_order.p, s.d_bptr.p, s.d_bcol.p, s.d_bval.p,
                               s.d_force.p, s.d_u.p, s.d_v.p, s.d_a.p, s.saved == SAVE_PLAIN ? nullptr : s.d_bc.p,
                               s.d_res.p, s.d_upd.p, st));
  s.prepared = -1;
  if (res_norm) {
    DFM_HIP(launch_dot(s.d_res.p, s.d_res.p, m, s.d_partial.p, s.d_counter.p, s.d_scal.p + S_TMP, st));
    DFM_HIP(hipMemcpyAsync(s.h_pinned_d, s.d_scal.p + S_TMP, sizeof(double), hipMemcpyDeviceToHost, st));
  }
  DFM_HIP(hipStreamSynchronize(st));
  if (res_norm) *res_norm = std::sqrt(s.h_pinned_d[0]);
  return DFM_OK;
}

int dfm_sys_update_field_save(dfm_sys sys) {
  NEED(sys, "dfm_sys_update_field_save");
  Sys& s = sys->s;
  if (s.saved == SAVE_DIA_NEWMARK || s.saved == SAVE_DIA_NEWMARKBETA) {
    const int rc = ensure_acc(s);
    if (rc != DFM_OK) return rc;
    DFM_HIP(launch_update_field_savedia(s.n, s.len, s.saved, s.save_dt, s.save_gamma, s.save_beta, s.d_order.p,
                                        s.d_bc.p, s.d_upd.p, s.d_u.p, s.d_v.p, s.d_a.p, s.ctx->stream));
    DFM_HIP(hipStreamSynchronize(s.ctx->stream));
    return DFM_OK;
  }
  DFM_HIP(launch_update_field_save(s.n, s.len, s.d_order.p, s.d_bc.p, s.d_upd.p, s.d_u.p, s.ctx->stream));
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  return DFM_OK;
}

int dfm_sys_get_force(dfm_sys sys, double* f) {
  NEED(sys, "dfm_sys_get_force");
  NEED(f, "dfm_sys_get_force");
  Sys& s = sys->s;
  DFM_CHECK(s.saved, DFM_ERR_STATE, "dfm_sys_get_force: assemble with dfm_sys_assemble_save first");
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  std::vector<double> h((size_t)mdof(s));
  DFM_HIP(hipMemcpy(h.data(), s.d_force.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int i0 = 0; i0 < s.n; i0++)
    std::memcpy(&f[(size_t)s.order[i0] * s.len], &h[(size_t)i0 * s.len], sizeof(double) * s.len);
  return DFM_OK;
}

int dfm_sys_boundary_terms(dfm_sys sys, long long* nterm) {
  NEED(sys, "dfm_sys_boundary_terms");
  NEED(nterm, "dfm_sys_boundary_terms");
  *nterm = sys->s.saved ? sys->s.nbterm : 0;
  return DFM_OK;
}

int dfm_sys_resize_tmp(dfm_sys sys, int ntmp) {
  NEED(sys, "dfm_sys_resize_tmp");
  Sys& s = sys->s;
  DFM_CHECK(ntmp >= 0, DFM_ERR_ARG, "dfm_sys_resize_tmp: negative size");
  // like CLinearSystem::ReSizeTmpVecSolver, only growth changes anything (src/ls/linearsystem.cpp:208-244)
  while ((int)s.d_tmp.size() < ntmp) {
    auto* b = new DBuf<double>();
    DFM_HIP(b->alloc(mdof(s)));
    DFM_HIP(launch_fill(b->p, 0.0, mdof(s), s.ctx->stream));
    s.d_tmp.push_back(b);
  }
  return DFM_OK;
}

int dfm_sys_ntmp(dfm_sys sys, int* ntmp) {
  NEED(sys, "dfm_sys_ntmp");
  NEED(ntmp, "dfm_sys_ntmp");
  *ntmp = (int)sys->s.d_tmp.size();
  return DFM_OK;
}

int dfm_sys_dot(dfm_sys sys, int iv1, int iv2, double* out) {
  NEED(sys, "dfm_sys_dot");
  NEED(out, "dfm_sys_dot");
  Sys& s = sys->s;
  double* a = vec_ptr(s, iv1);
  double* b = vec_ptr(s, iv2);
  DFM_CHECK(a && b, DFM_ERR_ARG, "dfm_sys_dot: bad vector handle");
  hipStream_t st = s.ctx->stream;
  DFM_HIP(launch_dot(a, b, mdof(s), s.d_partial.p, s.d_counter.p, s.d_scal.p + S_TMP, st));
  DFM_HIP(hipMemcpyAsync(s.h_pinned_d, s.d_scal.p + S_TMP, sizeof(double), hipMemcpyDeviceToHost, st));
  DFM_HIP(hipStreamSynchronize(st));
  *out = s.h_pinned_d[0];
  return DFM_OK;
}

int dfm_sys_copy(dfm_sys sys, int iv1, int iv2) {
  NEED(sys, "dfm_sys_copy");
  Sys& s = sys->s;
  double* a = vec_ptr(s, iv1);
  double* b = vec_ptr(s, iv2);
  DFM_CHECK(a && b, DFM_ERR_ARG, "dfm_sys_copy: bad vector handle");
  if (iv1 == iv2) return DFM_OK;
  DFM_HIP(launch_copy(a, b, mdof(s), s.ctx->stream));
  return DFM_OK;
}

int dfm_sys_scal(dfm_sys sys, double alpha, int iv) {
  NEED(sys, "dfm_sys_scal");
  Sys& s = sys->s;
  double* a = vec_ptr(s, iv);
  DFM_CHECK(a, DFM_ERR_ARG, "dfm_sys_scal: bad vector handle");
  DFM_HIP(launch_scal(a, alpha, mdof(s), s.ctx->stream));
  return DFM_OK;
}

int dfm_sys_axpy(dfm_sys sys, double alpha, int iv1, int iv2) {
  NEED(sys, "dfm_sys_axpy");
  Sys& s = sys->s;
  double* a = vec_ptr(s, iv1);
  double* b = vec_ptr(s, iv2);
  DFM_CHECK(a && b, DFM_ERR_ARG, "dfm_sys_axpy: bad vector handle");
  DFM_CHECK(iv1 != iv2, DFM_ERR_ARG, "dfm_sys_axpy: iv1 == iv2 is not supported (src/ls/linearsystem.cpp:318-321)");
  if (alpha == 0) return DFM_OK;
  DFM_HIP(launch_axpy(alpha, a, b, mdof(s), s.ctx->stream));
  return DFM_OK;
}

int dfm_sys_matvec(dfm_sys sys, double alpha, int iv1, double beta, int iv2) {
  NEED(sys, "dfm_sys_matvec");
  Sys& s = sys->s;
  double* x = vec_ptr(s, iv1);
  double* y = vec_ptr(s, iv2);
  DFM_CHECK(x && y, DFM_ERR_ARG, "dfm_sys_matvec: bad vector handle");
  if (alpha == 0.0) return dfm_sys_scal(sys, beta, iv2);
  DFM_CHECK(iv1 != iv2, DFM_ERR_ARG, "dfm_sys_matvec: iv1 == iv2 is not supported (src/ls/linearsystem.cpp:275-278)");
  DFM_CHECK(s.assembled, DFM_ERR_STATE, "dfm_sys_matvec: matrix not assembled");
  DFM_HIP(launch_spmv(s.n, s.len, spmv_tpr(s.ncrs, s.n), s.d_rowptr.p, s.d_colind.p, s.d_crs.p, s.d_dia.p, alpha, x,
                      beta, y, s.ctx->stream));
  return DFM_OK;
}

}  // extern "C"

// A vector handle into the caller's numbering, read on stream st (the stream whose work last wrote it).
int dfm::sys_download(Sys& s, int iv, double* out, hipStream_t st) {
  double* a = vec_ptr(s, iv);
  DFM_CHECK(a, DFM_ERR_ARG, "dfm_sys_get_vector: bad vector handle");
  std::vector<double> v((size_t)mdof(s));
  DFM_HIP(hipMemcpyAsync(v.data(), a, v.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  DFM_HIP(hipStreamSynchronize(st));
  for (int i0 = 0; i0 < s.n; i0++)
    for (int d = 0; d < s.len; d++) out[(size_t)s.order[i0] * s.len + d] = v[(size_t)i0 * s.len + d];
  return DFM_OK;
}

extern "C" {

int dfm_sys_get_vector(dfm_sys sys, int iv, double* out) {
  NEED(sys, "dfm_sys_get_vector");
  NEED(out, "dfm_sys_get_vector");
  return sys_download(sys->s, iv, out, sys->s.ctx->stream);
}

int dfm_sys_set_vector(dfm_sys sys, int iv, const double* in) {
  NEED(sys, "dfm_sys_set_vector");
  NEED(in, "dfm_sys_set_vector");
  Sys& s = sys->s;
  double* a = vec_ptr(s, iv);
  DFM_CHECK(a, DFM_ERR_ARG, "dfm_sys_set_vector: bad vector handle");
  std::vector<double> v((size_t)mdof(s));
  for (int i0 = 0; i0 < s.n; i0++)
    for (int d = 0; d < s.len; d++) v[(size_t)i0 * s.len + d] = in[(size_t)s.order[i0] * s.len + d];
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  s.prepared = -1;
  DFM_HIP(hipMemcpy(a, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  return DFM_OK;
}

// SetValue_Initialize + DoILUDecomp on the device, enqueued on st: the factor pattern takes the matrix values
// (ILU(0): the patterns coincide and the lane-parallel kernel reads the matrix itself; otherwise a scatter into
// the zeroed fill pattern, factored in place), then the phases of the numeric factorisation; singular pivots
// are counted in d_iflag[4].
// prefilled: the factor arrays already hold the values to factor (a multi-segment system's diagonal segment after
// its coupling updates); they are factored in place.
static hipError_t factor_numeric(Sys& s, hipStream_t st, bool prefilled = false) {
  hipError_t e;
  const long long fncrs = (long long)s.fcolind.size();
  const bool lanes = s.fplan.ok && std::getenv("DFM_ILU_ROWKERNEL") == nullptr;
  const double* a_crs = s.d_fcrs.p;
  const double* a_dia = s.d_fdia.p;
  if (prefilled) {
    // in place
  } else if (s.a2f.empty()) {
    if (lanes) a_crs = s.d_crs.p, a_dia = s.d_dia.p;
    else if ((e = launch_copy(s.d_crs.p, s.d_fcrs.p, (long long)s.ncrs * s.bs, st)) != hipSuccess) return e;
  } else {
    if ((e = launch_fill(s.d_fcrs.p, 0.0, fncrs * s.bs, st)) != hipSuccess) return e;
    if ((e = launch_scatter_blocks(s.ncrs, s.bs, s.d_a2f.p, s.d_crs.p, s.d_fcrs.p, st)) != hipSuccess) return e;
  }
  if (!prefilled && a_dia == s.d_fdia.p &&
      (e = launch_copy(s.d_dia.p, s.d_fdia.p, (long long)s.n * s.bs, st)) != hipSuccess)
    return e;
  if ((e = hipMemsetAsync(s.d_iflag.p + 4, 0, sizeof(int), st)) != hipSuccess) return e;
  if (lanes) {
    const FactorPlan& fp = s.fplan;
    FactorPlanDev pd{fp.d_gpass.p, fp.d_pmeta.p, fp.d_etab.p, fp.d_utab.p, fp.d_ukj.p, fp.d_ustep.p};
    return launch_ilu_rows(s.len, pd, s.fwd.phase_gptr.data(), s.fwd.nphase(), a_crs, a_dia, s.d_fcrs.p, s.d_fdia.p,
                           s.d_iflag.p + 4, st);
  }
  SweepDev fw{s.fwd.nphase(), s.fwd.phase_gptr.data(), s.fwd.phase_nlev.data(), s.fwd.d_grp.p,
              s.fwd.rowlist.empty() ? nullptr : s.fwd.d_rowlist.p, s.fwd.d_lev.p};
  return launch_ilu_factor(s.n, s.len, s.d_frowptr.p, s.d_fcolind.p, s.d_fdiaind.p, s.d_fcrs.p, s.d_fdia.p, fw,
                           s.d_iflag.p + 4, st);
}

}  // extern "C"

// The factor arrays, padded by two doubles as the matrix.
static int factor_alloc(Sys& s) {
  const long long fncrs = (long long)s.fcolind.size();
  if (s.d_fcrs.n != (size_t)(fncrs * s.bs + 2)) DFM_HIP(s.d_fcrs.alloc((size_t)(fncrs * s.bs + 2)));
  if (!s.d_fdia.p) DFM_HIP(s.d_fdia.alloc((size_t)s.n * s.bs + 2));
  return DFM_OK;
}

int dfm::sys_factor_prefill(Sys& s) {
  DFM_CHECK(s.assembled, DFM_ERR_STATE, "factor: matrix not assembled");
  int rc = factor_alloc(s);
  if (rc != DFM_OK) return rc;
  hipStream_t st = s.ctx->stream;
  const long long fncrs = (long long)s.fcolind.size();
  if (s.a2f.empty()) {
    DFM_HIP(launch_copy(s.d_crs.p, s.d_fcrs.p, (long long)s.ncrs * s.bs, st));
  } else {
    DFM_HIP(launch_fill(s.d_fcrs.p, 0.0, fncrs * s.bs, st));
    DFM_HIP(launch_scatter_blocks(s.ncrs, s.bs, s.d_a2f.p, s.d_crs.p, s.d_fcrs.p, st));
  }
  DFM_HIP(launch_copy(s.d_dia.p, s.d_fdia.p, (long long)s.n * s.bs, st));
  s.factored = s.factored32 = false;
  s.prepared = -1;
  return DFM_OK;
}

// The numeric factorisation (in place when prefilled), its sweep slabs and the singular-pivot check.
int dfm::sys_factor_run(Sys& s, bool prefilled) {
  hipStream_t st = s.ctx->stream;
  int rc = factor_alloc(s);
  if (rc != DFM_OK) return rc;
  DFM_HIP(factor_numeric(s, st, prefilled));
  int nsing = 0;
  DFM_HIP(hipMemcpyAsync(s.h_pinned + 8, s.d_iflag.p + 4, sizeof(int), hipMemcpyDeviceToHost, st));
  DFM_HIP(hipStreamSynchronize(st));
  nsing = s.h_pinned[8];
  s.factored32 = false;
  s.prepared = -1;
  // slab copies of the factor for the sweep kernels
  for (int dir = 0; dir < 2; dir++) {
    Sweep& w = dir == 0 ? s.fwd : s.bwd;
    const long long ne = w.nent;
    if (w.d_eval.n < (size_t)(ne * s.bs)) DFM_HIP(w.d_eval.alloc((size_t)std::max<long long>(ne, 1) * s.bs));
    DFM_HIP(launch_sweep_fill(s.len, (int)(w.gmeta.size() / 4), w.d_gmeta.p, w.d_scnt.p, w.d_esrc.p, s.d_fcrs.p,
                              w.d_eval.p, st));
    if (dir == 0) {
      const long long ns = (long long)w.srow.size();
      if (w.d_dval.n < (size_t)(ns * s.bs)) DFM_HIP(w.d_dval.alloc((size_t)ns * s.bs));
      DFM_HIP(launch_slab_fill(s.len, ns, (const unsigned*)w.d_srow.p, s.d_fdia.p, w.d_dval.p, st));
    }
  }
  if (nsing > 10) {
    s.factored = false;
    set_error("dfm_prec_factor: more than 10 singular pivots (DoILUDecomp returns false)");
    return DFM_ERR_SINGULAR;
  }
  s.factored = true;
  return DFM_OK;
}

extern "C" {

int dfm_prec_factor(dfm_sys sys) {
  NEED(sys, "dfm_prec_factor");
  Sys& s = sys->s;
  DFM_CHECK(s.assembled, DFM_ERR_STATE, "dfm_prec_factor: matrix not assembled");
  return sys_factor_run(s, false);
}

int dfm_prec_apply(dfm_sys sys, int iv) {
  NEED(sys, "dfm_prec_apply");
  Sys& s = sys->s;
  double* v = vec_ptr(s, iv);
  DFM_CHECK(v, DFM_ERR_ARG, "dfm_prec_apply: bad vector handle");
  DFM_CHECK(s.factored, DFM_ERR_STATE, "dfm_prec_apply: preconditioner not factored");
  DFM_HIP(precond_apply64(s, v, v, nullptr, s.ctx->stream));
  return DFM_OK;
}

int dfm_prec_set_fill_level(dfm_sys sys, int lev) {
  NEED(sys, "dfm_prec_set_fill_level");
  Sys& s = sys->s;
  DFM_CHECK(lev >= -1, DFM_ERR_ARG, "dfm_prec_set_fill_level: level must be >= -1");
  DFM_CHECK(!s.batch, DFM_ERR_STATE, "dfm_prec_set_fill_level: the system is being solved in a batch");
  if (lev == s.fill_level) return DFM_OK;
  const int old = s.fill_level;
  s.fill_level = lev;
  try {
    build_sweeps(s);
  } catch (const std::exception& ex) {
    s.fill_level = old;
    build_sweeps(s);
    set_error(std::string("dfm_prec_set_fill_level: ") + ex.what());
    return DFM_ERR_ARG;
  }
  // the matrix (pattern, values, SELL copy) is unchanged; the factor and its slabs must be computed again
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  s.factored = s.factored32 = false;
  s.prepared = -1;
  return upload_sweeps(s);
}

int dfm_prec_factor_size(dfm_sys sys, long long* fncrs) {
  NEED(sys, "dfm_prec_factor_size");
  NEED(fncrs, "dfm_prec_factor_size");
  *fncrs = (long long)sys->s.fcolind.size();
  return DFM_OK;
}

int dfm_prec_get_factor(dfm_sys sys, unsigned* frowptr, unsigned* fcolind, unsigned* diaind, double* fcrs,
                        double* fdia) {
  NEED(sys, "dfm_prec_get_factor");
  Sys& s = sys->s;
  DFM_CHECK(s.factored, DFM_ERR_STATE, "dfm_prec_get_factor: preconditioner not factored");
  DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  const size_t fncrs = s.fcolind.size();
  if (frowptr) std::memcpy(frowptr, s.frowptr.data(), sizeof(unsigned) * (s.n + 1));
  if (fcolind) std::memcpy(fcolind, s.fcolind.data(), sizeof(unsigned) * fncrs);
  if (diaind) std::memcpy(diaind, s.fdiaind.data(), sizeof(unsigned) * s.n);
  if (fcrs) DFM_HIP(hipMemcpy(fcrs, s.d_fcrs.p, sizeof(double) * fncrs * s.bs, hipMemcpyDeviceToHost));
  if (fdia) DFM_HIP(hipMemcpy(fdia, s.d_fdia.p, sizeof(double) * s.n * s.bs, hipMemcpyDeviceToHost));
  return DFM_OK;
}

static int ensure_work(Sys& s, int nvec64, int nvec32) {
  const long long m = mdof(s);
  if (s.d_work.n < (size_t)(nvec64 * m)) DFM_HIP(s.d_work.alloc((size_t)nvec64 * m));
  if (nvec32 && s.d_work32.n < (size_t)(nvec32 * m)) DFM_HIP(s.d_work32.alloc((size_t)nvec32 * m));
  return DFM_OK;
}

static int ensure_fp32(Sys& s) {
  if (s.factored32) return DFM_OK;
  hipStream_t st = s.ctx->stream;
  if (!s.d_crs32.p) DFM_HIP(s.d_crs32.alloc((size_t)s.ncrs * s.bs));
  if (!s.d_dia32.p) DFM_HIP(s.d_dia32.alloc((size_t)s.n * s.bs));
  DFM_HIP(launch_to_f32(s.d_crs.p, s.d_crs32.p, (long long)s.ncrs * s.bs, st));
  DFM_HIP(launch_to_f32(s.d_dia.p, s.d_dia32.p, (long long)s.n * s.bs, st));
  {
    const long long ne = s.sell.nslab * 64;
    if (s.sell.d_val32.n < (size_t)(ne * s.bs)) DFM_HIP(s.sell.d_val32.alloc((size_t)std::max<long long>(ne, 1) * s.bs));
    DFM_HIP(launch_slab_fill(s.len, ne, s.sell.d_src.p, s.d_crs.p, s.sell.d_val32.p, st));
  }
  for (int dir = 0; dir < 2; dir++) {
    Sweep& w = dir == 0 ? s.fwd : s.bwd;
    const long long ne = w.nent;
    if (w.d_eval32.n < (size_t)(ne * s.bs)) DFM_HIP(w.d_eval32.alloc((size_t)std::max<long long>(ne, 1) * s.bs));
    DFM_HIP(launch_sweep_fill(s.len, (int)(w.gmeta.size() / 4), w.d_gmeta.p, w.d_scnt.p, w.d_esrc.p, s.d_fcrs.p,
                              w.d_eval32.p, st));
    if (dir == 0) {
      const long long ns = (long long)w.srow.size();
      if (w.d_dval32.n < (size_t)(ns * s.bs)) DFM_HIP(w.d_dval32.alloc((size_t)ns * s.bs));
      DFM_HIP(launch_slab_fill(s.len, ns, (const unsigned*)w.d_srow.p, s.d_fdia.p, w.d_dval32.p, st));
    }
  }
  s.factored32 = true;
  return DFM_OK;
}

}  // extern "C"

namespace dfm {
int resolve_precision(const Sys& s, int precision) {
  if (precision == DFM_MIXED_AUTO) return s.len == 1 ? DFM_MIXED : DFM_FP64_TRUE;
  return precision;
}

int sys_solve_setup(Sys& s, int precision) {
  precision = resolve_precision(s, precision);
  DFM_CHECK(precision == DFM_FP64 || precision == DFM_MIXED || precision == DFM_FP32_ILU || precision == DFM_FP64_TRUE,
            DFM_ERR_ARG, "solve: bad precision");
  DFM_CHECK(s.assembled && s.factored, DFM_ERR_STATE, "solve: assemble and factor first");
  int rc = ensure_work(s, 2, precision == DFM_MIXED ? 5 : 0);
  if (rc != DFM_OK) return rc;
  if (precision == DFM_MIXED || precision == DFM_FP32_ILU) {
    rc = ensure_fp32(s);
    if (rc != DFM_OK) return rc;
  }
  DFM_HIP(launch_copy(s.d_res.p, s.d_rhs.p, mdof(s), s.ctx->stream));
  s.prepared = precision;
  return DFM_OK;
}
}  // namespace dfm

extern "C" {

int dfm_solve_pcg(dfm_sys sys, int precision, double* conv_ratio, unsigned* iteration) {
  NEED(sys, "dfm_solve_pcg");
  NEED(conv_ratio, "dfm_solve_pcg");
  NEED(iteration, "dfm_solve_pcg");
  Sys& s = sys->s;
  DFM_CHECK(!s.batch, DFM_ERR_STATE, "dfm_solve_pcg: the system is being solved in a batch");
  precision = resolve_precision(s, precision);
  const int rc = sys_solve_setup(s, precision);
  if (rc != DFM_OK) return rc;
  s.prepared = -1;  // consumed by this solve
  return solve_pcg(s, precision, conv_ratio, iteration);
}

int dfm_solve_pbicgstab(dfm_sys sys, int precision, double* conv_ratio, unsigned* iteration) {
  NEED(sys, "dfm_solve_pbicgstab");
  NEED(conv_ratio, "dfm_solve_pbicgstab");
  NEED(iteration, "dfm_solve_pbicgstab");
  Sys& s = sys->s;
  DFM_CHECK(!s.batch, DFM_ERR_STATE, "dfm_solve_pbicgstab: the system is being solved in a batch");
  DFM_CHECK(precision >= DFM_FP64 && precision <= DFM_FP64_TRUE, DFM_ERR_ARG, "dfm_solve_pbicgstab: bad precision");
  DFM_CHECK(s.assembled && s.factored, DFM_ERR_STATE, "dfm_solve_pbicgstab: assemble and factor first");
  // the mixed modes run the FP64 loop with the FP32 factor (the FP32-vector loop exists for PCG only)
  const bool f32 = precision == DFM_FP32_ILU || precision == DFM_MIXED || precision == DFM_MIXED_AUTO;
  const long long m = mdof(s);
  hipStream_t st = s.ctx->stream;
  int rc = ensure_work(s, 8, 0);  // the loop's seven vectors and the solution kept over a restart
  if (rc != DFM_OK) return rc;
  if (f32) {
    rc = ensure_fp32(s);
    if (rc != DFM_OK) return rc;
  }
  const double tol = *conv_ratio;
  const unsigned maxit = *iteration;
  DFM_HIP(launch_copy(s.d_res.p, s.d_rhs.p, m, st));
  double conv = tol;
  unsigned it = maxit;
  rc = solve_pbicgstab(s, f32, &conv, &it);
  if (rc < 0) return rc;
  bool ok = rc == 1 && it < maxit;
  unsigned used = it;
  if (ok && precision != DFM_FP64) {
    // as in dfm_solve_pcg: the loop stops on its recursive residual; its true residual is checked and the loop
    // restarted on r = b - A x (target half the tolerance, relative to ||b||) until the true residual meets the
    // tolerance; a restart that does not lower it by a fifth ends the solve unconverged
    double* xs = s.d_work.p + 7 * m;
    double prev = 0.0;
    for (int restart = 0;; restart++) {
      double rel = 0.0;
      rc = sys_true_residual(s, st, &rel);
      if (rc != DFM_OK) return rc;
      conv = rel;
      if (rel <= tol) break;
      if (used + 1 >= maxit || restart == kMaxRestarts || (restart > 0 && rel > kRestartGain * prev)) {
        ok = false;
        break;
      }
      prev = rel;
      DFM_HIP(launch_copy(s.d_upd.p, xs, m, st));
      DFM_HIP(launch_copy(s.d_rhs.p, s.d_res.p, m, st));
      DFM_HIP(launch_spmv(s.n, s.len, spmv_tpr(s.ncrs, s.n), s.d_rowptr.p, s.d_colind.p, s.d_crs.p, s.d_dia.p, -1.0,
                          xs, 1.0, s.d_res.p, st));
      double c2 = std::min(1.0, 0.5 * tol / rel);
      unsigned it2 = maxit - used - 1;
      rc = solve_pbicgstab(s, f32, &c2, &it2);
      if (rc < 0) return rc;
      DFM_HIP(launch_axpy(1.0, xs, s.d_upd.p, m, st));
      used += it2 + 1;
      if (rc != 1 || it2 >= maxit) {
        rc = sys_true_residual(s, st, &conv);
        if (rc != DFM_OK) return rc;
        ok = false;
        break;
      }
    }
  }
  *conv_ratio = conv;
  if (ok) *iteration = used;
  else if (precision == DFM_FP64) *iteration = it;
  else *iteration = maxit;
  return ok || precision == DFM_FP64 ? 1 : 0;
}

}  // extern "C"

// ||b - A x|| / ||b|| in FP64 for b = the assembled right-hand side and x = the update handle, enqueued on st
// and waited for; one pass over the matrix, the system's partial-sum buffer and scalar slots as scratch.
int dfm::sys_true_residual(Sys& s, hipStream_t st, double* rel) {
  DFM_HIP(launch_true_residual(s.n, s.len, spmv_tpr(s.ncrs, s.n), s.d_rowptr.p, s.d_colind.p, s.d_crs.p, s.d_dia.p,
                               s.d_upd.p, s.d_rhs.p, s.d_partial.p, s.d_scal.p + S_TMP, st));
  DFM_HIP(hipMemcpyAsync(s.h_pinned_d, s.d_scal.p + S_TMP, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  DFM_HIP(hipStreamSynchronize(st));
  *rel = s.h_pinned_d[1] > 0 ? std::sqrt(s.h_pinned_d[0] / s.h_pinned_d[1]) : 0.0;
  return DFM_OK;
}

extern "C" {

int dfm_sys_true_residual(dfm_sys sys, double* rel) {
  NEED(sys, "dfm_sys_true_residual");
  NEED(rel, "dfm_sys_true_residual");
  return sys_true_residual(sys->s, sys->s.ctx->stream, rel);
}

}  // extern "C"

// Average device time of `reps` back-to-back enqueues of launch() on st (after three untimed ones), HIP events.
template <typename F>
static int time_launches(hipStream_t st, int reps, F launch, double* avg_ms) {
  for (int w = 0; w < 3; w++) DFM_HIP(launch());
  hipEvent_t e0, e1;
  DFM_HIP(hipEventCreate(&e0));
  DFM_HIP(hipEventCreate(&e1));
  DFM_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < reps; r++) DFM_HIP(launch());
  DFM_HIP(hipEventRecord(e1, st));
  DFM_HIP(hipEventSynchronize(e1));
  float ms = 0;
  DFM_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (avg_ms) *avg_ms = ms / reps;
  return DFM_OK;
}

extern "C" {

int dfm_sys_bench_kernel(dfm_sys sys, int which, int reps, double* avg_ms, double* bytes) {
  NEED(sys, "dfm_sys_bench_kernel");
  Sys& s = sys->s;
  DFM_CHECK(which >= 0 && which <= 8 && reps > 0, DFM_ERR_ARG, "dfm_sys_bench_kernel: bad arguments");
  hipStream_t st = s.ctx->stream;
  if (which == 6) {
    // MakeResidual of the saved stiffness: force, residual, update and the field once per dof, the order map,
    // the term ranges, and per term its value, its column and the field value it gathers
    DFM_CHECK(s.saved == SAVE_PLAIN, DFM_ERR_STATE, "dfm_sys_bench_kernel: assemble with dfm_sys_assemble_save first");
    auto launch = [&]() {
      return launch_make_residual(s.n, s.len, SAVE_PLAIN, 0.0, 0.0, s.d_order.p, s.d_bptr.p, s.d_bcol.p, s.d_bval.p,
                                  s.d_force.p, s.d_u.p, nullptr, nullptr, nullptr, s.d_res.p, s.d_upd.p, st);
    };
    const int rc = time_launches(st, reps, launch, avg_ms);
    if (rc != DFM_OK) return rc;
    const double m = (double)mdof(s);
    if (bytes) *bytes = m * (4 * 8.0 + 4.0) + 4.0 + 4.0 * s.n + (double)s.nbterm * (8.0 + 4.0 + 8.0);
    s.prepared = -1;
    return DFM_OK;
  }
  if (which == 8) {
    // the assembly kernels of the last dfm_sys_assemble[_resident] (element store and gathers, or the per-row
    // gather into the zeroed matrix), rewriting the same matrix and residual; algorithmic bytes: the matrix,
    // diagonal and residual written once; coordinates, connectivity, per-element parameters and the field read
    DFM_CHECK(s.assembled && s.last_asm && !s.saved, DFM_ERR_STATE, "dfm_sys_bench_kernel: assemble first");
    AsmArgs a = *s.last_asm;
    ElemStore es{};
    const bool store = elem_store(s, a.kind, false, es);
    const long long m = mdof(s);
    auto launch = [&]() -> hipError_t {
      hipError_t e;
      if (!store) {
        if ((e = launch_fill(s.d_crs.p, 0.0, (long long)s.ncrs * s.bs, st)) != hipSuccess) return e;
        if ((e = launch_fill(s.d_dia.p, 0.0, (long long)s.n * s.bs, st)) != hipSuccess) return e;
        if ((e = launch_fill(s.d_res.p, 0.0, m, st)) != hipSuccess) return e;
      }
      return launch_assemble(a, store ? &es : nullptr, st);
    };
    const int rc = time_launches(st, reps, launch, avg_ms);
    if (rc != DFM_OK) return rc;
    // the matrix must be treated again (Dirichlet rows and columns, copies): a new assembly is needed to solve
    s.assembled = s.factored = s.factored32 = false;
    s.prepared = -1;
    if (bytes)
      *bytes = ((double)s.ncrs + s.n) * s.bs * 8.0 + (double)m * 8.0 + (double)s.n * s.ndim * 8.0 +
               (double)s.nelem * s.npoel * 4.0 + (double)s.nelem * (a.pstride ? a.pstride : 0) * 8.0 +
               (a.u ? (double)m * 8.0 : 0.0);
    return DFM_OK;
  }
  if (which == 7) {
    // the numeric ILU factorisation (initial values and factor kernels, not the slab copies): it rewrites the
    // factor with the same values. Algorithmic bytes: the matrix read and the factor written once (FP64 blocks),
    // the factor's column indices and row pointers
    DFM_CHECK(s.assembled && s.factored, DFM_ERR_STATE, "dfm_sys_bench_kernel: assemble and factor first");
    const int rc = time_launches(st, reps, [&]() { return factor_numeric(s, st); }, avg_ms);
    if (rc != DFM_OK) return rc;
    const double fn = (double)s.fcolind.size();
    if (bytes) *bytes = 2.0 * (fn + s.n) * s.bs * 8.0 + 4.0 * fn + 4.0 * (s.n + 1);
    return DFM_OK;
  }
  DFM_CHECK(s.assembled && s.factored, DFM_ERR_STATE, "dfm_sys_bench_kernel: assemble and factor first");
  int rc = ensure_work(s, 2, 5);
  if (rc != DFM_OK) return rc;
  if (which == 1 || which == 3 || which == 4) {
    rc = ensure_fp32(s);
    if (rc != DFM_OK) return rc;
  }
  return bench_kernel(s, which, reps, avg_ms, bytes);
}

int dfm_device_bytes(long long* bytes) {
  NEED(bytes, "dfm_device_bytes");
  *bytes = device_bytes_in_use();
  return DFM_OK;
}

int dfm_sys_last_solve_info(dfm_sys sys, int* reliable_updates, int* fell_back) {
  NEED(sys, "dfm_sys_last_solve_info");
  if (reliable_updates) *reliable_updates = sys->s.last_nrel;
  if (fell_back) *fell_back = sys->s.last_fallback;
  return DFM_OK;
}

int dfm_sys_last_solve_stats(dfm_sys sys, double* ms_total, double* ms_per_iter, unsigned* inner_iters) {
  NEED(sys, "dfm_sys_last_solve_stats");
  if (ms_total) *ms_total = sys->s.last_ms;
  if (ms_per_iter) *ms_per_iter = sys->s.last_ms_iter;
  if (inner_iters) *inner_iters = sys->s.last_inner;
  return DFM_OK;
}

}  // extern "C"
