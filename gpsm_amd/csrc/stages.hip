// Parity entry points for single pipeline stages (include/gpsm.h):
// gpsm_filter, gpsm_candidate_edges, gpsm_scan_u32, gpsm_compact_bitmap.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "api_internal.h"

using namespace gpsm;

namespace {

constexpr int STOP = 1;  // hook asks run_batch to stop (not an error)

struct FilterOut {
  uint32_t* bitmaps;
  uint32_t* counts;
  int rc;
};

int filter_hook(void* user, gpsm_ctx* c, const uint32_t* bitmaps, const uint32_t* counts, uint32_t nslots,
                uint32_t W) {
  FilterOut* fo = reinterpret_cast<FilterOut*>(user);
  fo->rc = GPSM_OK;
  if (fo->bitmaps &&
      hipMemcpyAsync(fo->bitmaps, bitmaps, sizeof(uint32_t) * (size_t)nslots * W, hipMemcpyDeviceToHost,
                     c->stream) != hipSuccess)
    fo->rc = GPSM_EDEVICE;
  if (hipStreamSynchronize(c->stream) != hipSuccess) fo->rc = GPSM_EDEVICE;
  if (fo->counts) std::memcpy(fo->counts, counts, sizeof(uint32_t) * nslots);
  return STOP;
}

struct CetOut {
  const gpsm_graph* g;
  uint32_t* lists;
  uint64_t cap;
  uint64_t* table_off;
  uint32_t* cet_size;
  int rc;
};

int cet_hook(void* user, gpsm_ctx* c, const CetDesc* fdesc, uint32_t nf) {
  CetOut* co = reinterpret_cast<CetOut*>(user);
  co->rc = GPSM_OK;
  for (uint32_t e = 0; e < nf; ++e) co->cet_size[e] = (uint32_t)fdesc[e].total;
  if (!co->lists) return STOP;
  // tables in the reference layout [ids | offsets | lists], back to back
  std::vector<CetDesc> d(fdesc, fdesc + nf);
  uint64_t offw = 0, listw = 0, total = 0;
  for (uint32_t e = 0; e < nf; ++e) {
    d[e].off_base = offw;
    d[e].list_base = listw;
    d[e].list_cap = (uint32_t)d[e].total;
    offw += d[e].ncand + 1;
    listw += d[e].total;
    co->table_off[e] = total;
    total += 2ull * d[e].ncand + 1 + d[e].total;
  }
  if (total > co->cap) {
    co->rc = GPSM_EOVERFLOW;
    return STOP;
  }
  for (auto& x : d) x.total = 0;
  hipStream_t s = c->stream;
  if (c->bufs[B_JDESC].ensure(sizeof(CetDesc) * std::max<uint32_t>(nf, 1)) ||
      c->bufs[B_OFFSETS].ensure(sizeof(uint32_t) * std::max<uint64_t>(offw, 1)) ||
      c->bufs[B_LISTS].ensure(sizeof(uint32_t) * std::max<uint64_t>(listw, 1)) ||
      c->bufs[B_CHUNKS].ensure(sizeof(uint32_t) * (nf + 2))) {
    co->rc = GPSM_ENOMEM;
    return STOP;
  }
  CetDesc* dd = reinterpret_cast<CetDesc*>(c->bufs[B_JDESC].ptr);
  uint32_t* offs = reinterpret_cast<uint32_t*>(c->bufs[B_OFFSETS].ptr);
  uint32_t* lists = reinterpret_cast<uint32_t*>(c->bufs[B_LISTS].ptr);
  const uint32_t* clist = reinterpret_cast<const uint32_t*>(c->bufs[B_CLIST].ptr);
  const uint32_t* bitmaps = reinterpret_cast<const uint32_t*>(c->bufs[B_BITMAPS].ptr);
  const DevGraph& dg = co->g->dg;
  const uint32_t grid = std::max<uint32_t>(c->ncu * 8, 64);
  hipMemcpyAsync(dd, d.data(), sizeof(CetDesc) * nf, hipMemcpyHostToDevice, s);
  launch_cet_count(dg, bitmaps, clist, dd, nf, offs, 1, reinterpret_cast<uint32_t*>(c->bufs[B_CHUNKS].ptr), grid, s);
  launch_cet_offsets(dd, nf, offs, s);
  launch_cet_write(dg, bitmaps, clist, dd, nf, offs, lists, reinterpret_cast<uint32_t*>(c->bufs[B_CHUNKS].ptr), grid,
                   s);
  std::vector<uint32_t> h_off(std::max<uint64_t>(offw, 1)), h_list(std::max<uint64_t>(listw, 1));
  hipMemcpyAsync(h_off.data(), offs, sizeof(uint32_t) * offw, hipMemcpyDeviceToHost, s);
  if (listw) hipMemcpyAsync(h_list.data(), lists, sizeof(uint32_t) * listw, hipMemcpyDeviceToHost, s);
  std::vector<uint32_t> h_clist;
  uint64_t cmax = 0;
  for (auto& x : d) cmax = std::max<uint64_t>(cmax, x.clist_off + x.ncand);
  h_clist.resize(std::max<uint64_t>(cmax, 1));
  hipMemcpyAsync(h_clist.data(), clist, sizeof(uint32_t) * cmax, hipMemcpyDeviceToHost, s);
  if (hipStreamSynchronize(s) != hipSuccess) {
    co->rc = GPSM_EDEVICE;
    return STOP;
  }
  for (uint32_t e = 0; e < nf; ++e) {
    uint32_t* t = co->lists + co->table_off[e];
    const uint32_t m = d[e].ncand;
    std::memcpy(t, h_clist.data() + d[e].clist_off, sizeof(uint32_t) * m);
    std::memcpy(t + m, h_off.data() + d[e].off_base, sizeof(uint32_t) * (m + 1));
    std::memcpy(t + 2 * m + 1, h_list.data() + d[e].list_base, sizeof(uint32_t) * fdesc[e].total);
  }
  return STOP;
}

}  // namespace

extern "C" {

int gpsm_filter(gpsm_ctx* c, const gpsm_graph* g, const gpsm_query* q, uint32_t* bitmaps, uint32_t* counts) {
  if (!c || !g || !q) return GPSM_EINVAL;
  HostQuery hq;
  int vr = hq.build(q);
  if (vr) return vr;
  FilterOut fo{bitmaps, counts, GPSM_OK};
  BatchHooks h;
  h.user = &fo;
  h.after_filter = filter_hook;
  gpsm_result r;
  int rc = run_batch(c, g, 1, q, nullptr, &r, &h);
  if (rc == STOP) return fo.rc;
  if (rc) return rc;
  return r.status ? r.status : GPSM_EINVAL;
}

int gpsm_candidate_edges(gpsm_ctx* c, const gpsm_graph* g, const gpsm_query* q, uint32_t* edge_from,
                         uint32_t* edge_to, uint32_t* cet_size, uint32_t* lists, uint64_t list_cap,
                         uint64_t* table_off) {
  if (!c || !g || !q || !edge_from || !edge_to || !cet_size || (lists && !table_off)) return GPSM_EINVAL;
  HostQuery hq;
  int rc = hq.build(q);
  if (rc) return rc;
  for (size_t e = 0; e < hq.ef.size(); ++e) {
    edge_from[e] = hq.ef[e];
    edge_to[e] = hq.et[e];
    cet_size[e] = 0;
  }
  CetOut co{g, lists, list_cap, table_off, cet_size, GPSM_OK};
  BatchHooks h;
  h.user = &co;
  h.after_cet_count = cet_hook;
  gpsm_result r;
  rc = run_batch(c, g, 1, q, nullptr, &r, &h);
  if (rc == STOP) return co.rc;
  if (rc) return rc;
  // the pipeline ended before the tables (an empty candidate set or a label
  // absent from the data graph): every table is empty
  return r.status;
}

int gpsm_scan_u32(gpsm_ctx* c, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* total) {
  if (!c || (n && (!in || !out))) return GPSM_EINVAL;
  if (total) *total = 0;
  if (n == 0) return GPSM_OK;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const uint64_t sw = scan_scratch_words(n) + 2;
  if (c->bufs[B_OFFSETS].ensure(sizeof(uint32_t) * n) || c->bufs[B_LISTS].ensure(sizeof(uint32_t) * n) ||
      c->bufs[B_SCRATCH].ensure(sizeof(uint32_t) * sw))
    return GPSM_ENOMEM;
  uint32_t* din = reinterpret_cast<uint32_t*>(c->bufs[B_OFFSETS].ptr);
  uint32_t* dout = reinterpret_cast<uint32_t*>(c->bufs[B_LISTS].ptr);
  uint32_t* scr = reinterpret_cast<uint32_t*>(c->bufs[B_SCRATCH].ptr);
  unsigned long long* dtot = reinterpret_cast<unsigned long long*>(scr);
  if (hipMemcpyAsync(din, in, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s) != hipSuccess) return GPSM_EDEVICE;
  launch_exclusive_scan(din, dout, n, dtot, scr + 2, sw - 2, s);
  unsigned long long t = 0;
  if (hipMemcpyAsync(out, dout, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&t, dtot, sizeof(t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPSM_EDEVICE;
  if (total) *total = t;
  return GPSM_OK;
}

int gpsm_compact_bitmap(gpsm_ctx* c, const uint32_t* bits, uint64_t nbits, uint32_t* ids, uint64_t* count) {
  if (!c || !count || (nbits && (!bits || !ids)) || nbits > 0xFFFFFFFFULL) return GPSM_EINVAL;
  *count = 0;
  if (nbits == 0) return GPSM_OK;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const uint64_t W = (nbits + 31) / 32;
  std::vector<uint32_t> h(bits, bits + W);
  if (nbits & 31) h[W - 1] &= (1u << (nbits & 31)) - 1u;  // bits past nbits are not part of the input
  const uint64_t sw = (compact_scratch_words(nbits) + 1) & ~1ull;  // dcount after it, 8-byte aligned
  if (c->bufs[B_BITMAPS].ensure(sizeof(uint32_t) * W) || c->bufs[B_CLIST].ensure(sizeof(uint32_t) * nbits) ||
      c->bufs[B_SCRATCH].ensure(sizeof(uint32_t) * (sw + 2)))
    return GPSM_ENOMEM;
  uint32_t* db = reinterpret_cast<uint32_t*>(c->bufs[B_BITMAPS].ptr);
  uint32_t* dids = reinterpret_cast<uint32_t*>(c->bufs[B_CLIST].ptr);
  uint32_t* scr = reinterpret_cast<uint32_t*>(c->bufs[B_SCRATCH].ptr);
  unsigned long long* dcount = reinterpret_cast<unsigned long long*>(scr + sw);
  if (hipMemcpyAsync(db, h.data(), sizeof(uint32_t) * W, hipMemcpyHostToDevice, s) != hipSuccess)
    return GPSM_EDEVICE;
  launch_compact_bitmap(db, nbits, dids, dcount, scr, s);
  unsigned long long cnt = 0;
  if (hipMemcpyAsync(&cnt, dcount, sizeof(cnt), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPSM_EDEVICE;
  if (cnt && hipMemcpy(ids, dids, sizeof(uint32_t) * cnt, hipMemcpyDeviceToHost) != hipSuccess) return GPSM_EDEVICE;
  *count = cnt;
  return GPSM_OK;
}

}  // extern "C"

extern "C" int gpsm_plan_features(const gpsm_graph* g, const gpsm_query* q, const uint32_t* cand,
                                  const uint64_t* cet_size, const uint32_t* order, const uint64_t* level_rows,
                                  float* feat, float* target) {
  if (!g || !q || !cand || !cet_size || !order || !level_rows || !feat || !target) return GPSM_EINVAL;
  HostQuery hq;
  int rc = hq.build(q);
  if (rc) return rc;
  std::vector<uint64_t> cs(cet_size, cet_size + hq.ef.size());
  OrderPlan pl;
  if (!order_from_vertices(hq, cs, order, pl)) return GPSM_EINVAL;
  std::vector<uint32_t> cv(cand, cand + hq.n);
  PlanContext pc{&hq, &g->stats, &cv, &cs};
  std::vector<uint8_t> placed(hq.n, 0);
  placed[order[0]] = placed[order[1]] = 1;
  int rows = 0;
  for (uint32_t k = 2; k < hq.n; ++k) {
    step_features(pc, placed, k, pl.par_edge[k], order[k], std::log1p((double)level_rows[k - 1]),
                  feat + (size_t)rows * GPSM_FEAT);
    target[rows] = (float)std::log1p((double)level_rows[k]);
    placed[order[k]] = 1;
    ++rows;
  }
  return rows;
}
