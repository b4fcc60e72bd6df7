// batch.hip -- many numeric form factors (own mesh, own q-grid, own rotation each) in one launch.
//
// The batch is cut into work units (item, q-tile of kWG x QPT q-points, chunk of triangle blocks) of
// similar cost, sorted longest-first and dispatched as one grid, so the hardware dispatcher hands
// the next unit to whichever CU frees up (longest-processing-time-first list scheduling). Uneven
// meshes (500 - 50 000 triangles) and grids (128^2 - 1024^2) then load-balance in a single launch.
// Units write disjoint FP64 partial slots; one finishing launch reduces them in chunk order
// (deterministic) and applies the per-q epilogue. Both launches can be replayed from a hipGraph.
// The reference computes one shape per call with its own allocations and copies
// (src/ff/gpu/ff_tri_gpu.cu:395-521).
#include <algorithm>

#include "numeric_core.hpp"

namespace hgx {

struct BatchItemDev {
  QView g;
  Rot R;
  Epi e;
  const float* rec32;
  const double* borig;
  const double* rec64;
  int64_t nblocks;
  int64_t ntri;
  int64_t bchunk;  // blocks per chunk
  int nchunks;
  double* part;
  double* out;
  float* part32;  // matrix-core items: FP32 partial sums instead of part
};

struct BatchUnit {
  int item;
  int qtile;
  int chunk;
  int pad;
};

struct FinishUnit {
  int item;
  int qblock;
};

template <int QPT>
__device__ inline void store_unit(const BatchItemDev& it, int chunk, int64_t qbase, const double (&ar)[QPT],
                                  const double (&ai)[QPT]) {
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    int64_t q = qbase + (int64_t)k * kWG;
    if (q >= it.g.nq) continue;
    // raw sums; k_batch_finish applies the epilogue (in place when there is a single chunk)
    double* dst = it.nchunks == 1 ? it.out + 2 * q : it.part + 2 * ((int64_t)chunk * it.g.nq + q);
    dst[0] = ar[k];
    dst[1] = ai[k];
  }
}

template <int QPT, bool CQ, bool KAHAN>
__global__ __launch_bounds__(kWG) void k_batch32(const BatchItemDev* __restrict__ items,
                                                 const BatchUnit* __restrict__ units) {
  const BatchUnit u = units[blockIdx.x];
  const BatchItemDev& it = items[u.item];
  const int64_t qbase = (int64_t)u.qtile * (kWG * QPT) + threadIdx.x;
  const int64_t b0 = (int64_t)u.chunk * it.bchunk;
  const int64_t b1 = min(it.nblocks, b0 + it.bchunk);
  double ar[QPT], ai[QPT];
  approx32_blocks<QPT, CQ, KAHAN>(it.g, it.R, qbase, it.rec32, it.borig, b0, b1, ar, ai);
  store_unit<QPT>(it, u.chunk, qbase, ar, ai);
}

template <int QPT, bool CQ>
__global__ __launch_bounds__(kWG) void k_batch64(const BatchItemDev* __restrict__ items,
                                                 const BatchUnit* __restrict__ units) {
  const BatchUnit u = units[blockIdx.x];
  const BatchItemDev& it = items[u.item];
  const int64_t qbase = (int64_t)u.qtile * (kWG * QPT) + threadIdx.x;
  const int64_t t0 = (int64_t)u.chunk * it.bchunk * kBlockTri;
  const int64_t t1 = min(it.ntri, t0 + it.bchunk * kBlockTri);
  cd mq[QPT][3];
  double ar[QPT], ai[QPT];
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    int64_t q = qbase + (int64_t)k * kWG;
    load_q<CQ>(it.g, it.R, q < it.g.nq ? q : it.g.nq - 1, mq[k]);
  }
  approx64_tris<QPT, CQ>(mq, it.rec64, t0, t1, ar, ai);
  store_unit<QPT>(it, u.chunk, qbase, ar, ai);
}

__global__ __launch_bounds__(kWG) void k_batch_finish(const BatchItemDev* __restrict__ items,
                                                      const FinishUnit* __restrict__ units) {
  const FinishUnit u = units[blockIdx.x];
  const BatchItemDev& it = items[u.item];
  const int64_t q = (int64_t)u.qblock * kWG + threadIdx.x;
  if (q >= it.g.nq) return;
  cd mq[3];
  load_q(it.g, it.R, q, mq);
  const cd raw = it.nchunks == 1 ? cd{it.out[2 * q], it.out[2 * q + 1]}
                                  : sum_chunks(it.nchunks, it.part, it.part32, it.g.nq, q);
  cd F = finish_q(it.e, mq, raw);
  it.out[2 * q] = F.r;
  it.out[2 * q + 1] = F.i;
}

}  // namespace hgx

using namespace hgx;

namespace {

using Batch = BatchState;

std::vector<Batch>& batches(Context* c) { return c->batches; }

template <int QPT>
void launch_batch(const Batch& b, hipStream_t s) {
  if (b.precision == HGX_PREC_FP64) {
    if (b.cq) k_batch64<QPT, true><<<b.nunits, kWG, 0, s>>>(static_cast<const BatchItemDev*>(b.d_items), static_cast<const BatchUnit*>(b.d_units));
    else k_batch64<QPT, false><<<b.nunits, kWG, 0, s>>>(static_cast<const BatchItemDev*>(b.d_items), static_cast<const BatchUnit*>(b.d_units));
  } else {
    bool kahan = (b.flags & HGX_FLAG_ACC_KAHAN) != 0;
    if (b.cq) {
      if (kahan) k_batch32<QPT, true, true><<<b.nunits, kWG, 0, s>>>(static_cast<const BatchItemDev*>(b.d_items), static_cast<const BatchUnit*>(b.d_units));
      else k_batch32<QPT, true, false><<<b.nunits, kWG, 0, s>>>(static_cast<const BatchItemDev*>(b.d_items), static_cast<const BatchUnit*>(b.d_units));
    } else {
      if (kahan) k_batch32<QPT, false, true><<<b.nunits, kWG, 0, s>>>(static_cast<const BatchItemDev*>(b.d_items), static_cast<const BatchUnit*>(b.d_units));
      else k_batch32<QPT, false, false><<<b.nunits, kWG, 0, s>>>(static_cast<const BatchItemDev*>(b.d_items), static_cast<const BatchUnit*>(b.d_units));
    }
  }
}

void launch_main(const Batch& b, hipStream_t s) {
  if (b.npunits > 0)
    product_launch_units(s, static_cast<const ProductItem*>(b.d_pitems), static_cast<const ProductUnit*>(b.d_punits),
                         b.npunits, b.cq, b.pmode);
  for (int k = 0; k < 4; ++k)  // Ewald-grid items, one launch per series order
    ewald_launch_units(s, static_cast<const EwaldItem*>(b.d_eitems),
                       static_cast<const EwaldUnit*>(b.d_eunits) + b.eunit_off[k], b.neunits[k], b.cq, k, b.epipe);
  if (b.nunits == 0) return;
  switch (b.qpt) {
    case 1: launch_batch<1>(b, s); break;
    case 2: launch_batch<2>(b, s); break;
    case 4: launch_batch<4>(b, s); break;
    default: launch_batch<8>(b, s); break;
  }
}

void launch_finish(const Batch& b, hipStream_t s) {
  if (b.nfinish > 0) k_batch_finish<<<b.nfinish, kWG, 0, s>>>(static_cast<const BatchItemDev*>(b.d_items), static_cast<const FinishUnit*>(b.d_finish));
}

}  // namespace

namespace hgx {
void free_batch_state(BatchState& b) {
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  if (b.d_items) (void)hipFree(b.d_items);
  if (b.d_units) (void)hipFree(b.d_units);
  if (b.d_finish) (void)hipFree(b.d_finish);
  if (b.d_part) (void)hipFree(b.d_part);
  if (b.d_pitems) (void)hipFree(b.d_pitems);
  if (b.d_punits) (void)hipFree(b.d_punits);
  if (b.d_prot) (void)hipFree(b.d_prot);
  if (b.d_eitems) (void)hipFree(b.d_eitems);
  if (b.d_eunits) (void)hipFree(b.d_eunits);
  b = BatchState();
}
}  // namespace hgx

extern "C" int hgx_tune_batch(hgx_ctx* h, int qpt, int blocks_per_unit) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if ((qpt != 1 && qpt != 2 && qpt != 4 && qpt != 8) || blocks_per_unit < 1) return HGX_ERR_INVALID_ARG;
  c->tune.batch_qpt = qpt;
  c->tune.batch_bpu = blocks_per_unit;
  return HGX_OK;
}

extern "C" int hgx_batch_create(hgx_ctx* h, int n, const hgx_batch_item* items, int precision, int flags, int* batch_id) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (n < 1 || !items || !batch_id) return set_error(c, HGX_ERR_INVALID_ARG, "batch_create: need items");
  if (precision != HGX_PREC_FP64 && precision != HGX_PREC_FP32C)
    return set_error(c, HGX_ERR_INVALID_ARG, "batch_create: bad precision");
  if (int rc = set_device(c)) return rc;
  Batch b;
  b.precision = precision;
  b.flags = flags;
  b.qpt = c->tune.batch_qpt;
  b.pmode = precision == HGX_PREC_FP64 ? product_mode64(c) : product_mode(c);
  b.epipe = c->tune.ewald_pipe != 0;
  const int64_t bpu = c->tune.batch_bpu;
  std::vector<BatchItemDev> it(n);
  std::vector<ProductItem> pit(n);
  std::vector<EwaldItem> eit(n);
  std::vector<char> is_prod(n, 0);
  std::vector<int> eord(n, -1);  // series orders of Ewald-grid items, -1 for the others
  std::vector<int64_t> part_off(n, -1), rot_off(n, -1);
  std::vector<char> f32part(n, 0);
  int64_t part_total = 0, rot_total = 0;  // doubles, bytes
  for (int i = 0; i < n; ++i) {
    const hgx_batch_item& x = items[i];
    Context* gc = x.grid ? reinterpret_cast<Context*>(x.grid) : c;
    if (!gc->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "batch_create: item %d has no q-grid", i);
    if (gc->device != c->device) return set_error(c, HGX_ERR_INVALID_ARG, "batch_create: item %d on another device", i);
    if (x.mesh_id < 0 || x.mesh_id >= (int)c->meshes.size() || !c->meshes[x.mesh_id].live)
      return set_error(c, HGX_ERR_BAD_HANDLE, "batch_create: item %d: unknown mesh %d", i, x.mesh_id);
    if (!x.ff) return set_error(c, HGX_ERR_INVALID_ARG, "batch_create: item %d has no output", i);
    const Mesh& m = c->meshes[x.mesh_id];
    b.grids.emplace_back(gc, gc->grid_gen);
    b.meshes.emplace_back(x.mesh_id, m.serial);
    BatchItemDev& d = it[i];
    d.g = QView{gc->d_qx, gc->d_qy, gc->d_qz, gc->npix, gc->nq};
    d.R = make_rot(x.rot);
    d.e.tv = make_vec(nullptr);
    d.e.has_tv = 0;
    d.e.volume = m.volume;
    d.e.moments = m.d_moments;
    for (int k = 0; k < 3; ++k) d.e.origin[k] = m.origin[k];
    d.e.radius = m.radius;
    d.e.use_series = precision == HGX_PREC_FP32C ? 1 : 0;
    d.rec32 = m.d_rec32;
    d.borig = m.d_borig;
    d.rec64 = m.d_rec64;
    d.nblocks = m.nblocks;
    d.ntri = m.ntri;
    d.bchunk = std::min<int64_t>(bpu, m.nblocks);
    const bool prod = gc->product && !(flags & HGX_FLAG_NO_MFMA);
    // Ewald-sphere grid (QGrid::create, detector), FP32C: matrix products with a short series
    // (ewald_mfma.hip), as hgx_ff_numeric takes them
    if (!prod && precision == HGX_PREC_FP32C && gc->ew.ok && !(flags & (HGX_FLAG_NO_MFMA | HGX_FLAG_ACC_KAHAN)))
      eord[i] = ewald_orders(gc, m.radius);
    if (eord[i] >= 0) {
      d.e.series_qr = kSeriesQRSplit;
      d.bchunk = std::min<int64_t>(product_blocks_per_unit(c), m.nblocks);
      EwaldItem& ei = eit[i];
      ewald_fill_item(gc, m, d.R, ei);
      ei.bchunk = d.bchunk;
      bool identity = true;
      for (int k = 0; k < 9; ++k) identity = identity && x.rot[k] == ((k % 4 == 0) ? 1.0 : 0.0);
      if (identity) {
        ei.rec = m.d_rec32;
        ei.borig = m.d_borig;
      } else {
        rot_off[i] = rot_total;
        rot_total += (int64_t)m.nblocks * (kBlockTri * 8 * sizeof(float) + 4 * sizeof(double));
      }
    }
    if (prod) {
      // product grid: complex matrix products on the matrix cores (product_mfma.hip)
      is_prod[i] = 1;
      d.e.series_qr = kSeriesQRSplit;
      d.bchunk = std::min<int64_t>(product_blocks_per_unit(c), m.nblocks);
      ProductItem& pi = pit[i];
      pi.qx = gc->d_qx;
      pi.qy = gc->d_qy;
      pi.qz = gc->d_qz;
      pi.nblocks = m.nblocks;
      pi.ntri = m.ntri;
      pi.bchunk = d.bchunk;
      pi.nq = gc->nq;
      pi.ncol = gc->ncol;
      pi.nz = gc->nbranch * gc->nrow;
      pi.snmax = (float)m.snmax;
      pi.rmax = (float)(std::sqrt(m.origin[0] * m.origin[0] + m.origin[1] * m.origin[1] + m.origin[2] * m.origin[2]) +
                        m.radius);
      bool identity = true;
      for (int k = 0; k < 9; ++k) identity = identity && x.rot[k] == ((k % 4 == 0) ? 1.0 : 0.0);
      if (identity) {
        pi.rec = m.d_rec32;
        pi.borig = m.d_borig;
        pi.rec64 = m.d_rec64;
      } else {
        rot_off[i] = rot_total;
        rot_total += precision == HGX_PREC_FP64 ? (int64_t)m.ntri * 8 * sizeof(double)
                                                : (int64_t)m.nblocks * (kBlockTri * 8 * sizeof(float) + 4 * sizeof(double));
      }
    }
    d.nchunks = (int)((m.nblocks + d.bchunk - 1) / d.bchunk);
    pit[i].nchunks = d.nchunks;
    eit[i].nchunks = d.nchunks;
    d.out = x.ff;
    d.part = nullptr;
    d.part32 = nullptr;
    // FP32C matrix-core items keep FP32 partial sums (half the doubles of space)
    f32part[i] = precision == HGX_PREC_FP32C && (prod || eord[i] >= 0);
    if (d.nchunks > 1) {
      part_off[i] = part_total;
      part_total += (int64_t)d.nchunks * gc->nq * (f32part[i] ? 1 : 2);
      part_total = (part_total + 31) & ~(int64_t)31;  // every item's partials 256-byte aligned
    }
    b.cq = b.cq || !gc->qz_real;
    b.pairs += (double)gc->nq * (double)m.ntri;
  }
  // work units, longest first
  struct U {
    double cost;
    BatchUnit u;
  };
  std::vector<U> us;
  struct PU {
    double cost;
    ProductUnit u;
  };
  std::vector<PU> pus;
  struct EU {
    double cost;
    EwaldUnit u;
  };
  std::vector<EU> eus[4];
  std::vector<FinishUnit> fin;
  const int64_t qtile = (int64_t)kWG * b.qpt;
  for (int i = 0; i < n; ++i) {
    const BatchItemDev& d = it[i];
    for (int64_t qb = 0; qb < (d.g.nq + kWG - 1) / kWG; ++qb) fin.push_back(FinishUnit{i, (int)qb});
    if (eord[i] >= 0) {
      const EwaldItem& ei = eit[i];
      const int nyg = ewald_col_groups(ei.ncol);
      // matrix work grows with the orders: split fp16 for orders 0 and 1, single fp16 above
      const double w = 3.0 * std::min(eord[i] + 1, 2) + std::max(eord[i] - 1, 0);
      for (int ch = 0; ch < d.nchunks; ++ch) {
        int64_t nb = std::min<int64_t>(d.bchunk, d.nblocks - ch * d.bchunk);
        for (int zt = 0; zt < ei.ntile; ++zt)
          for (int yg = 0; yg < nyg; ++yg) eus[eord[i]].push_back(EU{w * (double)nb, EwaldUnit{i, zt, yg, ch}});
      }
      continue;
    }
    if (is_prod[i]) {
      const ProductItem& pi = pit[i];
      int nzt = (int)ceil_div(pi.nz, product_zrows(b.pmode)), nyt = (int)ceil_div(pi.ncol, product_ycols(b.pmode));
      for (int ch = 0; ch < d.nchunks; ++ch) {
        int64_t nb = std::min<int64_t>(d.bchunk, d.nblocks - ch * d.bchunk);
        for (int zt = 0; zt < nzt; ++zt)
          for (int yt = 0; yt < nyt; ++yt) pus.push_back(PU{(double)nb, ProductUnit{i, zt, yt, ch}});
      }
      continue;
    }
    int64_t nqt = (d.g.nq + qtile - 1) / qtile;
    for (int64_t t = 0; t < nqt; ++t) {
      double nqpts = (double)std::min<int64_t>(qtile, d.g.nq - t * qtile);
      for (int ch = 0; ch < d.nchunks; ++ch) {
        int64_t nb = std::min<int64_t>(d.bchunk, d.nblocks - ch * d.bchunk);
        us.push_back(U{nqpts * (double)nb, BatchUnit{i, (int)t, ch, 0}});
      }
    }
  }
  std::stable_sort(us.begin(), us.end(), [](const U& a, const U& b2) { return a.cost > b2.cost; });
  std::stable_sort(pus.begin(), pus.end(), [](const PU& a, const PU& b2) { return a.cost > b2.cost; });
  std::vector<BatchUnit> units(us.size());
  for (size_t k = 0; k < us.size(); ++k) units[k] = us[k].u;
  std::vector<ProductUnit> punits(pus.size());
  for (size_t k = 0; k < pus.size(); ++k) punits[k] = pus[k].u;
  std::vector<EwaldUnit> eunits;
  for (int k = 0; k < 4; ++k) {
    std::stable_sort(eus[k].begin(), eus[k].end(), [](const EU& a, const EU& b2) { return a.cost > b2.cost; });
    b.eunit_off[k] = (int)eunits.size();
    b.neunits[k] = (int)eus[k].size();
    for (const EU& e : eus[k]) eunits.push_back(e.u);
  }
  if (units.size() > 0x7fffffff || punits.size() > 0x7fffffff || eunits.size() > 0x7fffffff)
    return set_error(c, HGX_ERR_INVALID_ARG, "batch_create: too many work units");
  b.nunits = (int)units.size();
  b.npunits = (int)punits.size();
  b.nfinish = (int)fin.size();
  auto fail = [&](const char* what) {
    (void)hipGetLastError();
    free_batch_state(b);
    return set_error(c, HGX_ERR_OOM, "batch_create: allocation of %s failed", what);
  };
  auto h2d = [](void* dst, const void* src, size_t n) {
    return hipMemcpy(dst, src, n, hipMemcpyHostToDevice) == hipSuccess;
  };
  auto copy_fail = [&]() {
    hipError_t e = hipGetLastError();
    free_batch_state(b);
    return set_error(c, HGX_ERR_DEVICE, "batch_create: copy to the device failed: %s", hipGetErrorString(e));
  };
  if (part_total > 0 && hipMalloc(&b.d_part, part_total * sizeof(double)) != hipSuccess) return fail("partials");
  for (int i = 0; i < n; ++i)
    if (part_off[i] >= 0 && f32part[i])
      it[i].part32 = pit[i].part32 = eit[i].part32 = reinterpret_cast<float*>(b.d_part + part_off[i]);
    else if (part_off[i] >= 0)
      it[i].part = pit[i].part = eit[i].part = b.d_part + part_off[i];
  for (int i = 0; i < n; ++i) pit[i].out = eit[i].out = it[i].out;
  if (rot_total > 0) {
    if (hipMalloc(&b.d_prot, rot_total) != hipSuccess) return fail("rotated records");
    for (int i = 0; i < n; ++i) {
      if (rot_off[i] < 0) continue;
      const Mesh& m = c->meshes[items[i].mesh_id];
      char* base = static_cast<char*>(b.d_prot) + rot_off[i];
      if (precision == HGX_PREC_FP64) {
        double* r64 = reinterpret_cast<double*>(base);
        if (int rc = product_rotate_records64(c, m, make_rot(items[i].rot), r64, c->stream)) {
          free_batch_state(b);
          return rc;
        }
        pit[i].rec64 = r64;
        continue;
      }
      float* rr = reinterpret_cast<float*>(base);
      double* ro = reinterpret_cast<double*>(base + (size_t)m.nblocks * kBlockTri * 8 * sizeof(float));
      if (int rc = product_rotate_mesh(c, m, make_rot(items[i].rot), rr, ro, c->stream)) {
        free_batch_state(b);
        return rc;
      }
      pit[i].rec = eit[i].rec = rr;
      pit[i].borig = eit[i].borig = ro;
    }
  }
  if (!eunits.empty()) {
    if (hipMalloc(&b.d_eitems, eit.size() * sizeof(EwaldItem)) != hipSuccess) return fail("Ewald-grid items");
    if (hipMalloc(&b.d_eunits, eunits.size() * sizeof(EwaldUnit)) != hipSuccess) return fail("Ewald-grid units");
    if (!h2d(b.d_eitems, eit.data(), eit.size() * sizeof(EwaldItem)) ||
        !h2d(b.d_eunits, eunits.data(), eunits.size() * sizeof(EwaldUnit)))
      return copy_fail();
  }
  if (b.npunits > 0) {
    if (hipMalloc(&b.d_pitems, pit.size() * sizeof(ProductItem)) != hipSuccess) return fail("product items");
    if (hipMalloc(&b.d_punits, punits.size() * sizeof(ProductUnit)) != hipSuccess) return fail("product units");
    if (!h2d(b.d_pitems, pit.data(), pit.size() * sizeof(ProductItem)) ||
        !h2d(b.d_punits, punits.data(), punits.size() * sizeof(ProductUnit)))
      return copy_fail();
  }
  if (hipMalloc(&b.d_items, it.size() * sizeof(BatchItemDev)) != hipSuccess) return fail("items");
  if (!units.empty() && hipMalloc(&b.d_units, units.size() * sizeof(BatchUnit)) != hipSuccess) return fail("units");
  if (!fin.empty() && hipMalloc(&b.d_finish, fin.size() * sizeof(FinishUnit)) != hipSuccess) return fail("finish");
  if (!h2d(b.d_items, it.data(), it.size() * sizeof(BatchItemDev)) ||
      (!units.empty() && !h2d(b.d_units, units.data(), units.size() * sizeof(BatchUnit))) ||
      (!fin.empty() && !h2d(b.d_finish, fin.data(), fin.size() * sizeof(FinishUnit))) ||
      hipStreamSynchronize(c->stream) != hipSuccess)  // rotated records ready before the first run
    return copy_fail();
  b.live = true;
  auto& B = batches(c);
  for (size_t i = 0; i < B.size(); ++i)
    if (!B[i].live) {
      B[i] = b;
      *batch_id = (int)i;
      return HGX_OK;
    }
  B.push_back(b);
  *batch_id = (int)B.size() - 1;
  return HGX_OK;
}

extern "C" int hgx_batch_run(hgx_ctx* h, int batch_id, int use_graph) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  auto& B = batches(c);
  if (batch_id < 0 || batch_id >= (int)B.size() || !B[batch_id].live)
    return set_error(c, HGX_ERR_BAD_HANDLE, "batch_run: unknown batch %d", batch_id);
  if (int rc = set_device(c)) return rc;
  Batch& b = B[batch_id];
  // the work tables hold the device addresses of the grids and meshes of creation time
  for (const auto& g : b.grids)
    if (g.first->grid_gen != g.second)
      return set_error(c, HGX_ERR_INVALID_ARG, "batch_run: a q-grid of batch %d was replaced or extended after "
                       "the batch was created; create the batch again", batch_id);
  for (const auto& m : b.meshes)
    if (m.first >= (int)c->meshes.size() || !c->meshes[m.first].live || c->meshes[m.first].serial != m.second)
      return set_error(c, HGX_ERR_INVALID_ARG, "batch_run: mesh %d of batch %d was released after the batch was "
                       "created", m.first, batch_id);
  if (!use_graph) {
    {
      LaunchTimer lt(c, true);
      launch_main(b, c->stream);
    }
    HGX_CHECK_LAUNCH(c);
    {
      LaunchTimer lt(c, false);
      launch_finish(b, c->stream);
    }
    HGX_CHECK_LAUNCH(c);
    return HGX_OK;
  }
  if (!b.exec) {
    // capture both launches once on a private stream, then replay on the context's stream
    hipStream_t cap;
    HGX_CHECK_HIP(c, hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    HGX_CHECK_HIP(c, hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    launch_main(b, cap);
    launch_finish(b, cap);
    hipError_t e1 = hipStreamEndCapture(cap, &b.graph);
    (void)hipStreamDestroy(cap);
    if (e1 != hipSuccess) return set_error(c, HGX_ERR_DEVICE, "batch_run: capture failed: %s", hipGetErrorString(e1));
    HGX_CHECK_HIP(c, hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
  }
  LaunchTimer lt(c, true);
  HGX_CHECK_HIP(c, hipGraphLaunch(b.exec, c->stream));
  return HGX_OK;
}

extern "C" int hgx_batch_info(hgx_ctx* h, int batch_id, int64_t* units, double* pairs) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  auto& B = batches(c);
  if (batch_id < 0 || batch_id >= (int)B.size() || !B[batch_id].live) return HGX_ERR_BAD_HANDLE;
  const Batch& bb = B[batch_id];
  if (units) *units = (int64_t)bb.nunits + bb.npunits + bb.neunits[0] + bb.neunits[1] + bb.neunits[2] + bb.neunits[3];
  if (pairs) *pairs = B[batch_id].pairs;
  return HGX_OK;
}

extern "C" int hgx_batch_release(hgx_ctx* h, int batch_id) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  auto& B = batches(c);
  if (batch_id < 0 || batch_id >= (int)B.size() || !B[batch_id].live)
    return set_error(c, HGX_ERR_BAD_HANDLE, "batch_release: unknown batch %d", batch_id);
  if (int rc = set_device(c)) return rc;
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  free_batch_state(B[batch_id]);
  return HGX_OK;
}
