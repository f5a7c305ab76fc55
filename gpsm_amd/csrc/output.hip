// gpsm_format_tuples / gpsm_write_tuples: the embedding lines of IO::output
// (io/IO.cpp:207-213) formatted on the device (kernels/text.hip), chunk by
// chunk so that device and host staging stay bounded whatever the result
// size.
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <memory>

#include "api_internal.h"
#include "gpsm.h"

using namespace gpsm;

namespace {

constexpr uint64_t kChunkOutBytes = 256ull << 20;  // device/host staging of one chunk's text

// Formats rows [0, n_rows) chunk by chunk; sink(text, bytes) receives every
// chunk's text in order and returns nonzero to stop.
template <typename Sink>
int format_chunks(gpsm_ctx* c, const uint32_t* tuples, uint64_t n_rows, uint32_t q, const uint8_t* keep,
                  uint64_t* bytes, Sink sink) {
  if (bytes) *bytes = 0;
  if (!n_rows) return GPSM_OK;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const uint32_t R = text_rows_per_wg(q);
  const uint64_t maxlen = text_fixed_length(q) + 10ull * q;
  // rows per chunk: a multiple of R whose text surely fits the staging
  const uint64_t ch = std::max<uint64_t>(R, (kChunkOutBytes / maxlen) / R * R);
  const uint64_t rows_cap = std::min(ch, n_rows);
  const uint64_t nwg_cap = (rows_cap + R - 1) / R;
  const uint64_t sw = scan_scratch_words(nwg_cap) + 2;
  const uint64_t out_cap = rows_cap * maxlen;
  // device buffers are the context's (reused across calls); the host
  // staging is ordinary memory
  if (c->bufs[B_TEXT_IN].ensure(sizeof(uint32_t) * rows_cap * q) ||
      c->bufs[B_TEXT_WG].ensure(sizeof(uint32_t) * (2 * nwg_cap + sw)) ||
      (keep && c->bufs[B_TEXT_KEEP].ensure(rows_cap)) || c->bufs[B_TEXT_OUT].ensure(out_cap))
    return GPSM_ENOMEM;
  uint32_t* dt = reinterpret_cast<uint32_t*>(c->bufs[B_TEXT_IN].ptr);
  uint32_t* scr = reinterpret_cast<uint32_t*>(c->bufs[B_TEXT_WG].ptr);  // 8-byte aligned: the total first
  uint32_t* dlen = scr + sw;
  uint32_t* doff = dlen + nwg_cap;
  uint8_t* dk = keep ? reinterpret_cast<uint8_t*>(c->bufs[B_TEXT_KEEP].ptr) : nullptr;
  char* dout = reinterpret_cast<char*>(c->bufs[B_TEXT_OUT].ptr);
  std::unique_ptr<char[]> hout(new char[out_cap]);
  unsigned long long* dtot = reinterpret_cast<unsigned long long*>(scr);
  uint64_t written = 0;
  int rc = GPSM_OK;
  for (uint64_t base = 0; base < n_rows && rc == GPSM_OK; base += ch) {
    const uint64_t rows = std::min(ch, n_rows - base);
    const uint64_t nwg = (rows + R - 1) / R;
    unsigned long long tot = 0;
    if (hipMemcpyAsync(dt, tuples + base * q, sizeof(uint32_t) * rows * q, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        (keep && hipMemcpyAsync(dk, keep + base, rows, hipMemcpyHostToDevice, s) != hipSuccess)) {
      rc = GPSM_EDEVICE;
      break;
    }
    launch_text_len(dt, rows, q, keep ? dk : nullptr, dlen, s);
    launch_exclusive_scan(dlen, doff, nwg, dtot, scr + 2, sw - 2, s);
    launch_text_write(dt, rows, q, keep ? dk : nullptr, doff, dout, s);
    if (hipMemcpyAsync(&tot, dtot, sizeof(tot), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) {
      rc = GPSM_EDEVICE;
      break;
    }
    if (tot && hipMemcpy(hout.get(), dout, tot, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = GPSM_EDEVICE;
      break;
    }
    rc = sink(hout.get(), tot, written);
    written += tot;
  }
  if (bytes) *bytes = written;
  return rc;
}

}  // namespace

extern "C" {

int gpsm_format_tuples(gpsm_ctx* c, const uint32_t* tuples, uint64_t n_rows, uint32_t q, const uint8_t* keep,
                       char* out, uint64_t cap, uint64_t* bytes) {
  if (!c || !q || q > GPSM_MAX_QV || (n_rows && !tuples) || (cap && !out)) return GPSM_EINVAL;
  bool over = false;
  const int rc = format_chunks(c, tuples, n_rows, q, keep, bytes, [&](const char* text, uint64_t n, uint64_t at) {
    if (at + n > cap) {
      over = true;
      if (at < cap) std::memcpy(out + at, text, cap - at);
    } else {
      std::memcpy(out + at, text, n);
    }
    return GPSM_OK;
  });
  if (rc != GPSM_OK) return rc;
  return over ? GPSM_EOVERFLOW : GPSM_OK;
}

int gpsm_write_tuples(gpsm_ctx* c, int fd, const uint32_t* tuples, uint64_t n_rows, uint32_t q, const uint8_t* keep,
                      uint64_t* bytes) {
  if (!c || fd < 0 || !q || q > GPSM_MAX_QV || (n_rows && !tuples)) return GPSM_EINVAL;
  return format_chunks(c, tuples, n_rows, q, keep, bytes, [&](const char* text, uint64_t n, uint64_t) {
    while (n) {
      const ssize_t w = ::write(fd, text, n);
      if (w < 0) {
        if (errno == EINTR) continue;
        return GPSM_EINVAL;
      }
      text += w;
      n -= (uint64_t)w;
    }
    return GPSM_OK;
  });
}

}  // extern "C"
