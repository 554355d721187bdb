// ASCII output of result records on the host cores: the printf formatting
// the reference programs do value by value (rcontrib's "%.6e\t" records,
// rt/rc2.c:306; rtrace's "%e\t", rt/rtrace.c:912-916; gendaymtx's
// "%.3g %.3g %.3g\n" rows, gen/gendaymtx.c:652-663), split over threads by
// record ranges.  Each
// value goes through snprintf, so the text is the C library's own.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/radgpu.h"

namespace {

// One conversion: '%', optional ".N" with N <= 17, then e or g (at most
// 25 characters for any double).
bool valid_format(const char* f) {
  if (!f || f[0] != '%') return false;
  const char* p = f + 1;
  if (*p == '.') {
    ++p;
    int prec = 0, nd = 0;
    while (*p >= '0' && *p <= '9' && nd < 2) prec = 10 * prec + (*p++ - '0'), ++nd;
    if (nd == 0 || prec > 17) return false;
  }
  if (*p != 'e' && *p != 'g') return false;
  return p[1] == '\0';
}

void format_range(const double* v, long long r0, long long r1, int per_record, long long group, const char* fmt,
                  const char* sep, const char* tail, std::string& s) {
  char buf[64];
  const size_t lsep = strlen(sep), ltail = strlen(tail);
  for (long long r = r0; r < r1; ++r) {
    const double* x = v + r * per_record;
    for (int k = 0; k < per_record; ++k) {
      int n = snprintf(buf, sizeof buf, fmt, x[k]);
      s.append(buf, (size_t)n);
      if (k + 1 < per_record) s.append(sep, lsep);
    }
    s.append(tail, ltail);
    s.push_back('\n');
    if (group > 0 && (r + 1) % group == 0) s.push_back('\n');
  }
}

}  // namespace

extern "C" {

long long rg_format_ascii(const double* values, long long nrecords, int per_record, long long group,
                          const char* fmt, const char* sep, const char* tail, char* out, long long cap, int nthreads) {
  if (nrecords < 0 || per_record <= 0 || !valid_format(fmt) || !sep || !tail || (nrecords && (!values || !out)))
    return -1;
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  nthreads = (int)std::min<long long>(nthreads, std::max<long long>(1, nrecords / 4096));
  // chunk boundaries fall on record boundaries; a group's blank line follows
  // its last record whichever chunk holds it
  std::vector<std::string> parts((size_t)nthreads);
  std::vector<std::thread> th;
  long long per = (nrecords + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    long long r0 = std::min(nrecords, per * t), r1 = std::min(nrecords, r0 + per);
    auto job = [=, &parts] {
      parts[(size_t)t].reserve((size_t)(r1 - r0) * (size_t)per_record * 14);
      format_range(values, r0, r1, per_record, group, fmt, sep, tail, parts[(size_t)t]);
    };
    if (t + 1 == nthreads)
      job();
    else
      th.emplace_back(job);
  }
  for (auto& x : th) x.join();
  long long total = 0;
  for (auto& p : parts) total += (long long)p.size();
  if (total > cap) return -2 - total;  // caller's buffer too small: -(2 + needed)
  long long off = 0;
  for (auto& p : parts) {
    memcpy(out + off, p.data(), p.size());
    off += (long long)p.size();
  }
  return total;
}

}  // extern "C"
