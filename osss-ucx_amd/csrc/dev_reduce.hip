// dev_reduce.hip - C-ABI entries of the team reduction (see dev_reduce.h):
// the per-call launch, the batched tile loop and the tile-span query.
#include "dev_reduce.h"

namespace osc {

int visit_reduce(int dtype, int op, int mode, PerCallV *pc, BatchV *bv,
                 SpanV *sv) {
  switch (dtype) {
    case OSC_T_U32: case OSC_T_U64: case OSC_T_I32: case OSC_T_I64:
      return visit_reduce_int_wide(dtype, op, mode, pc, bv, sv);
    case OSC_T_U8: case OSC_T_U16: case OSC_T_I8: case OSC_T_I16:
      return visit_reduce_int_narrow(dtype, op, mode, pc, bv, sv);
    case OSC_T_F32: return visit_reduce_f32(dtype, op, mode, pc, bv, sv);
    case OSC_T_F64: return visit_reduce_f64(dtype, op, mode, pc, bv, sv);
    case OSC_T_C32: case OSC_T_C64:
      return visit_reduce_cplx(dtype, op, mode, pc, bv, sv);
  }
  return OSC_ERR_ARG;
}

}  // namespace osc

#include <string.h>

#include <mutex>
#include <string>
#include <unordered_map>

namespace osc {
int g_split_group = -1;
bool g_split_partial = true;
bool g_split_walk_rule = true;
thread_local LaunchNote t_launch[kMaxLaunchNames];
thread_local int t_nlaunch = 0;
thread_local int t_split_phase = -1;
thread_local uint8_t *t_split_flags = nullptr;

// Span flags of split float MIN/MAX reductions: one device buffer per
// stream, since calls on different streams may run at the same time while
// calls on one stream run in order (the walk writes every flag of its call,
// the replicate that follows it on the stream reads them).  Allocated on a
// stream's first such call; never under graph capture (the call then runs
// as one kernel).
namespace {
std::mutex g_flag_mu;
std::unordered_map<hipStream_t, uint8_t *> g_flag_bufs;
uint64_t g_flag_cap = 0;                 // bytes per buffer
constexpr size_t kMaxFlagStreams = 256;  // beyond: one-kernel calls
}  // namespace

uint8_t *span_flags(hipStream_t s, uint64_t npieces) {
  std::lock_guard<std::mutex> lk(g_flag_mu);
  if (npieces == 0 || npieces > g_flag_cap) return nullptr;
  auto it = g_flag_bufs.find(s);
  if (it != g_flag_bufs.end()) return it->second;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess ||
      st != hipStreamCaptureStatusNone)
    return nullptr;
  if (g_flag_bufs.size() >= kMaxFlagStreams) return nullptr;
  void *p = nullptr;
  if (hipMalloc(&p, g_flag_cap) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  g_flag_bufs.emplace(s, (uint8_t *)p);
  return (uint8_t *)p;
}
// Replicate layout by measurement.  The fastest replicate layout -- members
// per workgroup G and workgroups per interleave group S -- differs between
// MI355X boxes.  With the branch-free visit body, one member per workgroup
// leads everywhere measured: at 64 members x 32 MiB (232 MiB apart) one box
// took 316 / 321 / 337 us with (1, 16) / (1, 32) / (1, 64), 347 with
// (2, 32) and 352-375 with 4 or 8 members; another took 308-310 us with
// (1, 32) and 313-315 with (1, 16) and (1, 64).  (With one member per
// workgroup the 63 groups of a super-chunk run together and workgroup x of
// each runs on the same XCD, so the 63 reads of a piece after the first
// come from that XCD's L2.)  So from 8 MiB per member, each power-of-two
// size class measures four candidates on its first calls, without ever
// waiting: each of its first eight calls (whatever their team size) runs
// the next candidate between two events (two rounds), and a later call that
// finds every pair of events complete fixes the candidate with the lowest
// time per byte written.  Below 8 MiB, until then, under graph capture and
// when the table is full, calls take (1, 32).
namespace {
constexpr uint64_t kTuneMinBytes = 8ull << 20;
constexpr int kTuneSlots = 64;
constexpr int kCand = 4, kRounds = 2;
constexpr int kCandG[kCand] = {1, 1, 1, 2};
constexpr int kCandS[kCand] = {16, 32, 64, 32};
// below kTuneMinBytes, under graph capture and while undecided
constexpr int kFixedG = 1, kFixedS = 32;
struct TuneSlot {
  int cls;
  int choice;    // candidate index, -1 while undecided
  int recorded;  // timed runs recorded, 0..kCand*kRounds
  hipEvent_t ev[2 * kCand * kRounds];
  double bytes[kCand * kRounds];  // bytes each timed run wrote
};
std::mutex g_tune_mu;
TuneSlot g_tune[kTuneSlots];
int g_tune_n = 0;

int size_class(uint64_t nbytes) { return 63 - __builtin_clzll(nbytes); }
}  // namespace

// The replicate kernel's name as rocprof prints it: k_replicate<G, S, SKIP>.
const char *replicate_name(int group, int width, bool skip) {
#define OSC_REP_NAMES(S, B)                                         \
  {"k_replicate<1, " #S ", " #B ">", "k_replicate<2, " #S ", " #B ">", \
   "k_replicate<4, " #S ", " #B ">", "k_replicate<8, " #S ", " #B ">", \
   "k_replicate<16, " #S ", " #B ">"}
  static const char *const names[2][3][5] = {
      {OSC_REP_NAMES(16, false), OSC_REP_NAMES(32, false),
       OSC_REP_NAMES(64, false)},
      {OSC_REP_NAMES(16, true), OSC_REP_NAMES(32, true),
       OSC_REP_NAMES(64, true)}};
#undef OSC_REP_NAMES
  const int i = group == 1 ? 0 : group == 2 ? 1 : group == 4 ? 2
              : group == 16 ? 4 : 3;
  const int j = width == 64 ? 2 : width == 32 ? 1 : 0;
  return names[skip ? 1 : 0][j][i];
}

int replicate_run(const oscd_geom *g, uint64_t off, uint64_t nbytes,
                  int group, int width, const uint8_t *skip, hipStream_t s) {
  if (group < 0) {
    group = kFixedG;
    width = kFixedS;
  }
  note_launch(replicate_name(group, width, skip != nullptr), nullptr);
  return oscd_replicate(g, off, nbytes, group, width, skip, s);
}

int replicate_tuned(const oscd_geom *g, uint64_t off, uint64_t nbytes,
                    const uint8_t *skip, hipStream_t s) {
  if (g_split_group > 0 || nbytes < kTuneMinBytes)
    return replicate_run(g, off, nbytes, g_split_group, 16, skip, s);
  const int cls = size_class(nbytes);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  TuneSlot *t = nullptr;
  for (int i = 0; i < g_tune_n; ++i)
    if (g_tune[i].cls == cls) t = &g_tune[i];
  if (t != nullptr && t->choice >= 0)
    return replicate_run(g, off, nbytes, kCandG[t->choice],
                         kCandS[t->choice], skip, s);
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess ||
      st != hipStreamCaptureStatusNone)
    return replicate_run(g, off, nbytes, kFixedG, kFixedS, skip, s);
  if (t == nullptr) {
    if (g_tune_n >= kTuneSlots) return replicate_run(g, off, nbytes, kFixedG, kFixedS, skip, s);
    TuneSlot nt{};
    nt.cls = cls;
    nt.choice = -1;
    for (int e = 0; e < 2 * kCand * kRounds; ++e)
      if (hipEventCreate(&nt.ev[e]) != hipSuccess) {
        for (int f = 0; f < e; ++f) (void)hipEventDestroy(nt.ev[f]);
        (void)hipGetLastError();
        return replicate_run(g, off, nbytes, kFixedG, kFixedS, skip, s);
      }
    g_tune[g_tune_n] = nt;
    t = &g_tune[g_tune_n++];
  }
  if (t->recorded < kCand * kRounds) {
    const int r = t->recorded, c = r % kCand;
    if (hipEventRecord(t->ev[2 * r], s) != hipSuccess) return OSC_ERR_HIP;
    const int rc = replicate_run(g, off, nbytes, kCandG[c], kCandS[c], skip, s);
    if (rc != OSC_OK) return rc;
    if (hipEventRecord(t->ev[2 * r + 1], s) != hipSuccess) return OSC_ERR_HIP;
    t->bytes[r] = (double)(g->npes - 1) * (double)nbytes;
    t->recorded = r + 1;
    return OSC_OK;
  }
  // every run recorded: decide once all of them have completed (calls of
  // one size class may come on different streams)
  bool done = true;
  for (int r = 0; r < kCand * kRounds && done; ++r)
    done = hipEventQuery(t->ev[2 * r + 1]) == hipSuccess;
  (void)hipGetLastError();
  if (done) {
    float best[kCand];
    for (int c = 0; c < kCand; ++c) best[c] = 1e30f;
    for (int r = 0; r < kCand * kRounds; ++r) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, t->ev[2 * r], t->ev[2 * r + 1]) ==
              hipSuccess &&
          t->bytes[r] > 0.0) {
        const float per_byte = (float)(ms / t->bytes[r]);
        if (per_byte < best[r % kCand]) best[r % kCand] = per_byte;
      }
    }
    int bc = 0;
    for (int c = 1; c < kCand; ++c)
      if (best[c] < best[bc]) bc = c;
    t->choice = bc;
    for (int e = 0; e < 2 * kCand * kRounds; ++e) {
      (void)hipEventDestroy(t->ev[e]);
      t->ev[e] = nullptr;
    }
    (void)hipGetLastError();
    return replicate_run(g, off, nbytes, kCandG[bc], kCandS[bc], skip, s);
  }
  return replicate_run(g, off, nbytes, kFixedG, kFixedS, skip, s);
}

int replicate_choice(int npes, uint64_t nbytes, int *width) {
  if (width != nullptr) *width = 16;
  if (g_split_group > 0) return g_split_group;
  if (width != nullptr) *width = kFixedS;
  if (nbytes < kTuneMinBytes) return kFixedG;
  if (width != nullptr) *width = 16;
  const int cls = size_class(nbytes);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  (void)npes;  // the size class alone keys the measurement
  for (int i = 0; i < g_tune_n; ++i)
    if (g_tune[i].cls == cls) {
      if (g_tune[i].choice < 0) return 0;
      if (width != nullptr) *width = kCandS[g_tune[i].choice];
      return kCandG[g_tune[i].choice];
    }
  return 0;
}

void replicate_tune_reset() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int i = 0; i < g_tune_n; ++i)
    for (int e = 0; e < 2 * kCand * kRounds; ++e)
      if (g_tune[i].ev[e] != nullptr) (void)hipEventDestroy(g_tune[i].ev[e]);
  g_tune_n = 0;
}
}  // namespace osc

extern "C" int oscd_replicate_group(int npes, uint64_t nbytes, int *width) {
  return osc::replicate_choice(npes, nbytes, width);
}

extern "C" void oscd_set_span_flag_capacity(uint64_t bytes_per_member) {
  std::lock_guard<std::mutex> lk(osc::g_flag_mu);
  osc::g_flag_cap = bytes_per_member / 4096 + 1;
}

extern "C" void oscd_release_span_flags(void) {
  {
    std::lock_guard<std::mutex> lk(osc::g_flag_mu);
    for (auto &kv : osc::g_flag_bufs) (void)hipFree(kv.second);
    osc::g_flag_bufs.clear();
    osc::g_flag_cap = 0;
  }
  osc::replicate_tune_reset();
}

extern "C" void oscd_set_split_phase(int phase) {
  osc::t_split_phase = phase == 0 || phase == 1 ? phase : -1;
}

extern "C" void oscd_set_split_flags(void *flags) {
  osc::t_split_flags = (uint8_t *)flags;
}

extern "C" void oscd_set_split_partial(int on) {
  osc::g_split_partial = on != 0;
}

extern "C" void oscd_set_split_walk_rule(int on) {
  osc::g_split_walk_rule = on != 0;
}

extern "C" void oscd_set_reduce_split(int enabled, int group) {
  osc::g_split_group =
      !enabled ? 0
      : (group == 1 || group == 2 || group == 4 || group == 8 || group == 16)
          ? group
          : -1;
}

static int log2_floor(int p) {
  int lg = 0;
  while ((2 << lg) <= p) ++lg;
  return lg;
}

extern "C" int oscd_reduce(const oscd_geom *g, int op, int dtype, int fpmode,
                           int shape, uint64_t dest_off, uint64_t src_off,
                           uint64_t n, void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || g->npes > 8191 || op < 0 ||
      op >= OSC_OP_COUNT)
    return OSC_ERR_ARG;
  osc::t_nlaunch = 0;
  if (n == 0) return OSC_OK;
  PerCallV v{g, log2_floor(g->npes), dest_off, src_off, n, shape,
             (hipStream_t)stream};
  return visit_reduce(dtype, op, fpmode, &v, nullptr, nullptr);
}

// "const char *osc::k_reduce_pf() [T = double, Pol = osc::PolF64Acc<5>,
// VEC = 2, ...]" -> "k_reduce<double, osc::PolF64Acc<5>, 2, ...>"
static std::string kernel_name(const osc::LaunchNote &nt) {
  if (nt.pretty == nullptr) return nt.what;
  const std::string pf = nt.pretty;
  const size_t a = pf.find('['), b = pf.rfind(']');
  if (a == std::string::npos || b == std::string::npos || b < a) return nt.what;
  std::string out = std::string(nt.what) + "<";
  int depth = 0;
  bool first = true;
  size_t i = a + 1, item = i;
  for (; i <= b; ++i) {
    const char c = pf[i];
    if (c == '<' || c == '(') ++depth;
    if (c == '>' || c == ')') --depth;
    if ((c == ',' && depth == 0) || i == b) {
      const std::string kv = pf.substr(item, i - item);
      const size_t eq = kv.find(" = ");
      if (!first) out += ", ";
      out += eq == std::string::npos ? kv : kv.substr(eq + 3);
      first = false;
      item = i + 2;  // skip ", "
    }
  }
  return out + ">";
}

extern "C" int oscd_last_reduce_kernels(char *buf, size_t len) {
  std::string all;
  for (int i = 0; i < osc::t_nlaunch; ++i) {
    if (i) all += ";";
    all += kernel_name(osc::t_launch[i]);
  }
  if (buf != nullptr && len > 0) {
    const size_t k = all.size() < len - 1 ? all.size() : len - 1;
    memcpy(buf, all.data(), k);
    buf[k] = '\0';
  }
  return osc::t_nlaunch;
}

extern "C" int oscd_batch_reduce(int dtype, int op, int fpmode, int shape,
                                 int maxlg, const oscd_bcall *calls,
                                 const oscd_tile *tiles, uint32_t ntiles,
                                 uint32_t grid, size_t lds_bytes, char *base,
                                 uint64_t slice, void *stream) {
  using namespace osc;
  if (grid == 0) grid = 1;
  BatchV v{{calls, tiles, ntiles, grid, lds_bytes, base, slice, maxlg,
            (hipStream_t)stream},
           shape};
  return visit_reduce(dtype, op, fpmode, nullptr, &v, nullptr);
}

extern "C" int oscd_reduce_span_units(int dtype, int op, int fpmode,
                                      int shape) {
  using namespace osc;
  SpanV v{shape};
  return visit_reduce(dtype, op, fpmode, nullptr, nullptr, &v);
}

extern "C" int oscd_reduce_vec(int dtype) {
  static const int vec[OSC_T_COUNT] = {8, 8, 4, 2, 8, 8, 4, 2, 4, 2, 2, 1};
  return (dtype < 0 || dtype >= OSC_T_COUNT) ? 0 : vec[dtype];
}
