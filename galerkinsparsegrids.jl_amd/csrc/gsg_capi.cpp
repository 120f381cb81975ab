// gsg_capi.cpp -- the C ABI of libgsg (include/gsg.h): operator objects, the
// device-resident RK4 driver, batched solves and the continuous-batching server.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <sstream>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/gsg.h"
#include "gsg_dev.hpp"
#include "gsg_setup.hpp"

namespace gsg {

// ---------------------------------------------------------------------------------
// errors

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

static thread_local std::string g_last_error;

static void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(GSG_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(x) hip_check((x), #x)

template <typename F>
static int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return GSG_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return GSG_ENOMEM;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return GSG_EINVAL;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return GSG_EINVAL;
  }
}

static void require(bool ok, const std::string& msg) {
  if (!ok) throw Error(GSG_EINVAL, msg);
}

// ---------------------------------------------------------------------------------
// device memory

// Stream capture is fragile against calls from other threads that synchronize the device
// (hipFree, hipMalloc, synchronous copies, stream and event creation or destruction): a
// capture running on one thread (an operator's RK steps, a batch, the server's worker)
// would be invalidated.  Every such call and every capture section holds this mutex.
static std::recursive_mutex g_capture_mu;
typedef std::lock_guard<std::recursive_mutex> CaptureGuard;

struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  explicit DevMem(size_t b) { alloc(b); }
  void alloc(size_t b) {
    release();
    if (b == 0) return;
    CaptureGuard cg(g_capture_mu);
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw Error(e == hipErrorOutOfMemory ? GSG_ENOMEM : GSG_EDEVICE,
                  std::string("hipMalloc(") + std::to_string(b) + "): " + hipGetErrorString(e));
    }
    bytes = b;
  }
  void release() {
    if (p) {
      CaptureGuard cg(g_capture_mu);
      (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
  }
  ~DevMem() { release(); }
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  DevMem(DevMem&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevMem& operator=(DevMem&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

template <typename T>
static void upload(DevMem& m, const std::vector<T>& v) {
  m.alloc(std::max<size_t>(v.size(), 1) * sizeof(T));
  CaptureGuard cg(g_capture_mu);
  if (!v.empty()) HIPC(hipMemcpy(m.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

static void ensure_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    throw Error(GSG_EDEVICE, "no HIP device available");
  }
}
// makes `dev` the calling thread's device for a scope and restores the previous one
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) HIPC(hipSetDevice(dev));
  }
  ~DeviceScope() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// ---------------------------------------------------------------------------------
// operator tables, cached per (D, k, n, scheme, device)

struct Tables {
  int D, k, n, full, device;
  Layout lt;
  Op1D op;
  CellPlan plan;
  bool tile_ok = false;     // the fiber kernels (LDS or direct) can serve this grid
  TilePlan tplan[2];        // per precision (index: fp32), sized to the LDS budget
  std::vector<Tile> ttiles; // tile-plan tiles (blk = tile index), natural order
  std::vector<Tile> dtiles; // direct-kernel tiles: one cell each, blk = packed list lengths,
                            // aux = device address of the cell's plan entries
  DevMem blk_lev, blk_off, blk_inoff, row_ptr, ent, val64, val32, val32p, plan_ptr, plan_ent;
  DevMem tile_bptr[2], batch_sptr[2], slot_off[2], slot_dir[2], batch_lptr[2], tile_ent[2];
  DevTables devp[2];        // device tables with the precision's tile plan
  static int64_t ipow_i(int64_t b, int e) {
    int64_t r = 1;
    while (e-- > 0) r *= b;
    return r;
  }
  std::vector<Tile> tiles;  // tiles over [0, N) for solve index 0
  // device memory the tables hold (every array, both precisions' tile plans)
  size_t device_bytes() const {
    size_t b = blk_lev.bytes + blk_off.bytes + blk_inoff.bytes + row_ptr.bytes + ent.bytes + val64.bytes +
               val32.bytes + val32p.bytes + plan_ptr.bytes + plan_ent.bytes;
    for (int f = 0; f < 2; ++f)
      b += tile_bptr[f].bytes + batch_sptr[f].bytes + slot_off[f].bytes + slot_dir[f].bytes + batch_lptr[f].bytes +
           tile_ent[f].bytes;
    return b;
  }
};

static std::mutex g_tab_mu;
static std::map<std::tuple<int, int, int, int, int, int>, std::shared_ptr<Tables>> g_tabs;

static std::vector<Tile> make_tiles(const Layout& lt, int solve) {
  std::vector<Tile> t;
  int blk = 0;
  for (int64_t s = 0; s < lt.N; s += TILE) {
    while (s >= lt.blk_off[blk + 1]) ++blk;
    t.push_back(Tile{solve, blk, s, std::min<int64_t>(s + TILE, lt.N), 0});
  }
  return t;
}

// Workgroups are dispatched to the 8 XCDs round robin (workgroup id mod 8), and each
// XCD has its own L2.  Interleave per-XCD tile lists so that bin x runs on XCD x; short
// bins are padded with empty tiles that repeat a real tile with an empty range (so the
// addresses a kernel reads before it sees the empty range stay valid).
constexpr int NXCD = 8;
// chunk: tiles per workgroup (the direct kernel runs DIRECT_CPW tiles per workgroup);
// bin x's chunks go to workgroups x, x + 8, ...
static std::vector<Tile> xcd_interleave(const std::vector<std::vector<Tile>>& bins, size_t chunk) {
  size_t L = 0;  // chunks per bin
  Tile pad{0, 0, 0, 0, 0};
  for (auto& b : bins) {
    L = std::max(L, (b.size() + chunk - 1) / chunk);
    if (!b.empty()) pad = Tile{b[0].solve, b[0].blk, b[0].start, b[0].start, b[0].aux};
  }
  std::vector<Tile> out(L * NXCD * chunk, pad);
  for (size_t x = 0; x < bins.size(); ++x)
    for (size_t j = 0; j < bins[x].size(); ++j) out[((j / chunk) * NXCD + x) * chunk + j % chunk] = bins[x][j];
  return out;
}
// contiguous eighths of a tile list, one per XCD, each starting at a multiple of `unit`
// tiles (the segmented kernel's wavefronts must not straddle two solves)
static std::vector<Tile> xcd_split(const std::vector<Tile>& tiles, size_t chunk = 1, size_t unit = 1) {
  std::vector<std::vector<Tile>> bins(NXCD);
  size_t per = (tiles.size() + NXCD - 1) / NXCD;
  per = (per + unit - 1) / unit * unit;
  for (size_t t = 0; t < tiles.size(); ++t) bins[per ? t / per : 0].push_back(tiles[t]);
  return xcd_interleave(bins, chunk);
}
// Visit order of single-cell tiles: by the Morton code of the cell centre.  A cell's
// coupled input cells overlap it in space (at every level), so a spatially coherent
// order keeps the inputs in flight on an XCD within a small region and lets its L2
// serve them; vector order (level block after level block) cycles through the whole
// state instead.  With group > 1 the tiles move in runs of `group` consecutive tiles
// (a segmented-kernel wavefront's cells), keyed by each run's first cell.
static void spatial_order(const Layout& lt, std::vector<Tile>& tiles, size_t group = 1) {
  const int D = lt.D;
  int64_t kD = 1;
  for (int e = 0; e < D; ++e) kD *= lt.k;
  const int bits = std::min(20, 60 / D);  // bits per dimension of the key
  if (tiles.size() % group) throw std::logic_error("tile runs must be whole");
  std::vector<std::pair<uint64_t, size_t>> key(tiles.size() / group);
  for (size_t t = 0; t < key.size(); ++t) {
    const int64_t g = tiles[t * group].start / kD;
    int b = int(std::upper_bound(lt.blk_off.begin(), lt.blk_off.end(), g * kD) - lt.blk_off.begin()) - 1;
    const int64_t cl = g - lt.blk_off[b] / kD;
    uint64_t code = 0;
    int cum = 0;
    for (int e = 0; e < D; ++e) {
      const int l = lt.blk_lev[b * D + e];
      const int sh = l > 0 ? l - 1 : 0;
      const uint64_t c = uint64_t(cl >> cum) & ((uint64_t(1) << sh) - 1);
      cum += sh;
      // centre (c + 1/2) 2^-sh in units of 2^-bits
      const uint64_t x = ((2 * c + 1) << bits) >> (sh + 1);
      for (int i = 0; i < bits; ++i) code |= ((x >> i) & 1) << (i * D + e);
    }
    key[t] = {code, t};
  }
  std::stable_sort(key.begin(), key.end());
  std::vector<Tile> out;
  out.reserve(tiles.size());
  for (size_t t = 0; t < key.size(); ++t)
    for (size_t q = 0; q < group; ++q) out.push_back(tiles[key[t].second * group + q]);
  tiles.swap(out);
}

// table kinds: the derivative operator, or a basis transform (its 1-D block table applied
// along one direction; the plan kernel only)
enum : int { TAB_DERIV = 0, TAB_M2N = 1, TAB_N2M = 2, TAB_N2P = 3, TAB_P2N = 4 };

static std::shared_ptr<Tables> get_tables(int D, int k, int n, int full, int tkind = TAB_DERIV) {
  require(D >= 1 && D <= MAXD, "D must be in 1..8");
  require(k >= 1 && k <= KMAX, "k must be in 1..10");
  require(n >= 0 && n <= 20, "n must be in 0..20");
  require(full == 0 || full == 1, "scheme must be GSG_SCHEME_SPARSE or GSG_SCHEME_FULL");
  ensure_device();
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  auto key = std::make_tuple(D, k, n, full, dev, tkind);
  {
    std::lock_guard<std::mutex> g(g_tab_mu);
    auto it = g_tabs.find(key);
    if (it != g_tabs.end()) return it->second;
  }
  auto T = std::make_shared<Tables>();
  T->D = D; T->k = k; T->n = n; T->full = full; T->device = dev;
  T->lt = build_layout(D, k, n, full != 0);
  // kernels address each vector (at most D*N elements of 8 bytes) with 31-bit byte offsets
  require(T->lt.N > 0 && int64_t(D) * T->lt.N * 8 < (int64_t(1) << 31),
          "grid too large: D * get_size * 8 bytes must stay below 2 GiB per vector");
  if (tkind == TAB_DERIV) {
    T->op = build_op1d(k, n);
  } else {
    require(nodal_k_ok(k), "the nodal basis exists for k in {2, 3, 5}");
    require(n <= 12, "basis transforms: n must be in 0..12");
    static const int ft[5][2] = {{0, 0}, {TB_MODAL, TB_NODAL}, {TB_NODAL, TB_MODAL}, {TB_NODAL, TB_POINTS}, {TB_POINTS, TB_NODAL}};
    T->op = transform_op1d(k, n, ft[tkind][0], ft[tkind][1]);
  }
  T->plan = build_cell_plan(T->lt, T->op);
  upload(T->blk_lev, T->lt.blk_lev);
  upload(T->blk_off, T->lt.blk_off);
  upload(T->blk_inoff, T->lt.blk_inoff);
  upload(T->row_ptr, T->op.row_ptr);
  upload(T->ent, T->op.ent);
  std::vector<double> v64(T->op.val);
  v64.resize(v64.size() + size_t(k) * k, 0.0);  // the zero block the plan pads with
  upload(T->val64, v64);
  // The float operator is the FP64 table split into two float tables, hi = fl(v) and
  // lo = fl(v - hi), stored one after the other: sum_j hi x_j + sum_j lo x_j, each
  // accumulated in float and added in double, carries the table to about 2^-48 relative,
  // so the only float roundings left are the stage input's and the accumulations', which
  // change from stage to stage and do not add up over a run the way a fixed table
  // rounding does.
  std::vector<float> v32(2 * v64.size());
  for (size_t i = 0; i < v64.size(); ++i) {
    v32[i] = float(v64[i]);
    v32[v64.size() + i] = float(v64[i] - double(v32[i]));
  }
  upload(T->val32, v32);
  std::vector<float> v32p;  // hi blocks, then lo blocks
  if (k <= 4) {  // the blocks padded to 4 x 4 for the segmented kernel's MFMA rows
    const size_t nb = v64.size() / (size_t(k) * k);
    v32p.assign(2 * nb * 16, 0.f);
    for (size_t e = 0; e < nb; ++e)
      for (int r = 0; r < k; ++r)
        for (int c = 0; c < k; ++c) {
          const size_t i = (e * k + r) * k + c;
          v32p[e * 16 + r * 4 + c] = v32[i];
          v32p[(nb + e) * 16 + r * 4 + c] = v32[v64.size() + i];
        }
    upload(T->val32p, v32p);
  }
  upload(T->plan_ptr, T->plan.ptr);
  {
    // 16 spare words: the direct kernel reads a cell's entries 8 at a time
    std::vector<int32_t> pe(T->plan.ent);
    pe.resize(pe.size() + 16, 0);
    upload(T->plan_ent, pe);
  }
  const int64_t kD = Tables::ipow_i(k, D);
  // the shapes with a fiber-kernel instantiation; the others run the plan kernel
  T->tile_ok = tkind == TAB_DERIV && fiber_ok(D, k);
  if (T->tile_ok && !direct_ok(D, k)) {  // the direct kernel needs no tile plan
    // LDS per workgroup: 32 KiB of input cells, the batch's coefficient blocks (at most
    // 2048 elements, none when they are wave-uniform) and the staging of offsets and
    // entries (at most 2048): below 64 KiB in both precisions
    for (int f = 0; f < 2; ++f) {
      const int es = f ? 4 : 8;
      const int ent_elems = fiber_uniform(D, k) ? 2048 * k * k : 2048;
      T->tplan[f] = build_tile_plan(T->lt, T->op, T->plan, fiber_cpt(D, k), 32768 / es, ent_elems, 2048, 1);
      upload(T->tile_bptr[f], T->tplan[f].tile_bptr);
      upload(T->batch_sptr[f], T->tplan[f].batch_sptr);
      upload(T->slot_off[f], T->tplan[f].slot_off);
      upload(T->slot_dir[f], T->tplan[f].slot_dir);
      upload(T->batch_lptr[f], T->tplan[f].batch_lptr);
      upload(T->tile_ent[f], T->tplan[f].ent);
    }
    if (T->tplan[0].ntiles != T->tplan[1].ntiles) throw std::logic_error("tile plans disagree");
    const int64_t ncell = T->lt.N / kD;
    for (int t = 0; t < T->tplan[0].ntiles; ++t) {
      const int64_t g0 = int64_t(t) * T->tplan[0].cpt;
      const int64_t g1 = std::min<int64_t>(g0 + T->tplan[0].cpt, ncell);
      T->ttiles.push_back(Tile{0, t, g0 * kD, g1 * kD, 0});
    }
  }
  for (int f = 0; f < 2; ++f)
    T->devp[f] = DevTables{T->blk_lev.as<int32_t>(),     T->blk_off.as<int64_t>(),     T->blk_inoff.as<int64_t>(),
                           T->row_ptr.as<int32_t>(),     T->ent.as<int32_t>(),         T->val64.as<double>(),
                           T->val32.as<float>(),         T->plan_ptr.as<int32_t>(),    T->plan_ent.as<int32_t>(),
                           T->tile_bptr[f].as<int32_t>(), T->batch_sptr[f].as<int32_t>(), T->slot_off[f].as<int32_t>(),
                           T->slot_dir[f].as<int32_t>(), T->batch_lptr[f].as<int32_t>(), T->tile_ent[f].as<int32_t>(),
                           int64_t(v64.size()), T->val32p.as<float>(), int64_t(v32p.size() / 2)};
  T->tiles = make_tiles(T->lt, 0);
  if (T->tile_ok && (direct_ok(D, k) || seg_ok(D, k))) {
    // one tile per cell; the D list lengths (< 256 each) packed a byte per direction
    const int64_t ncell = T->lt.N / kD;
    T->dtiles.reserve(size_t(ncell));
    for (int64_t g = 0; g < ncell; ++g) {
      uint32_t lens = 0;
      for (int e = 0; e < D; ++e) {
        const int32_t len = T->plan.ptr[g * D + e + 1] - T->plan.ptr[g * D + e];
        if (len > 255) throw std::logic_error("plan list too long for the direct kernel");
        lens |= uint32_t(len) << (8 * e);
      }
      const int64_t addr = int64_t(reinterpret_cast<uintptr_t>(T->plan_ent.as<int32_t>() + 2 * int64_t(T->plan.ptr[g * D])));
      T->dtiles.push_back(Tile{0, int32_t(lens), g * kD, (g + 1) * kD, addr});
    }
    if (!seg_ok(D, k) || seg_cpw(D, k) == 1) {
      spatial_order(T->lt, T->dtiles);
    } else {
      // the segmented kernel takes seg_cpw consecutive cells a wavefront: pad to whole
      // wavefronts with empty tiles (readable entry address, empty range)
      const Tile last = T->dtiles.back();
      while (T->dtiles.size() % size_t(seg_cpw(D, k))) T->dtiles.push_back(Tile{0, 0, last.end, last.end, last.aux});
      // wavefront runs visited in Morton order of their first cell (the runs stay whole)
      spatial_order(T->lt, T->dtiles, size_t(seg_cpw(D, k)));
    }
  }
  std::lock_guard<std::mutex> g(g_tab_mu);
  auto it = g_tabs.find(key);
  if (it != g_tabs.end()) return it->second;
  g_tabs[key] = T;
  return T;
}

// ---------------------------------------------------------------------------------
// kernel selection: built-in rule, tuning table, GSG_KERNEL environment override

static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int>, int> g_tune;
static std::once_flag g_tune_loaded;
// the launch schedule of FP32 batches and servers whose own setting is GSG_SCHEDULE_AUTO
// (a "schedule S" line of the tuning file, or gsg_tune_schedule); -1: not set
static int g_tune_schedule = -1;
// streams a batch spreads its classes over under the streams schedule (a "lanes L" line
// of the tuning file, or gsg_tune_lanes; 1..8); -1: not set (4)
static int g_tune_lanes = -1;

// The tuning file holds one line "D k n precision variant" per grid (precision 0 = FP64,
// 1 = FP32; variant GSG_KERNEL_*) and at most one line "schedule S" (S =
// GSG_SCHEDULE_STREAMS or GSG_SCHEDULE_FUSED); '#' starts a comment.  It is read once, from
// $GSG_TUNE_FILE if set, else from gsg_tune.txt beside the library; a missing file
// leaves the built-in rule.  Entries set by gsg_tune_set take precedence.
static void load_tune_file() {
  std::string path;
  if (const char* e = std::getenv("GSG_TUNE_FILE")) {
    path = e;
  } else {
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&gsg_version), &info) && info.dli_fname) {
      path = info.dli_fname;
      const size_t slash = path.find_last_of('/');
      path = (slash == std::string::npos ? std::string() : path.substr(0, slash + 1)) + "gsg_tune.txt";
    }
  }
  if (path.empty()) return;
  std::ifstream in(path);
  std::string line;
  std::lock_guard<std::mutex> g(g_tune_mu);
  while (std::getline(in, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream ls(line);
    if (line.find("lanes") != std::string::npos) {
      std::string word;
      int nl = 0;
      if ((ls >> word >> nl) && word == "lanes" && nl >= 1 && nl <= 8 && g_tune_lanes < 0) g_tune_lanes = nl;
      continue;
    }
    if (line.find("schedule") != std::string::npos) {
      std::string word;
      int sch = 0;
      if ((ls >> word >> sch) && word == "schedule" && (sch == GSG_SCHEDULE_STREAMS || sch == GSG_SCHEDULE_FUSED) &&
          g_tune_schedule < 0)
        g_tune_schedule = sch;
      continue;
    }
    int D, k, n, prec, var;
    if (!(ls >> D >> k >> n >> prec >> var)) continue;
    if (D < 1 || D > MAXD || k < 1 || k > KMAX || n < 0 || n > 20 || (prec != 0 && prec != 1)) continue;
    if (var != GSG_KERNEL_AUTO && var != GSG_KERNEL_PLAN && var != GSG_KERNEL_TILE && var != GSG_KERNEL_TILE_ALT)
      continue;
    g_tune.emplace(std::make_tuple(D, k, n, prec), var);  // keeps entries set before the load
  }
}

// streams a batch's classes spread over (GSG_LANES in the environment, else the tuned
// count, else 4)
static int tuned_lanes() {
  if (const char* e = std::getenv("GSG_LANES")) {
    const int nl = std::atoi(e);
    if (nl >= 1 && nl <= 8) return nl;
  }
  std::call_once(g_tune_loaded, load_tune_file);
  std::lock_guard<std::mutex> g(g_tune_mu);
  return g_tune_lanes > 0 ? g_tune_lanes : 4;
}

// the schedule a batch runs: its own setting, else GSG_SCHEDULE in the environment
// ("streams" / "fused"), else the tuned default, else streams
static int resolve_schedule(int own) {
  if (own == GSG_SCHEDULE_STREAMS || own == GSG_SCHEDULE_FUSED) return own;
  if (const char* e = std::getenv("GSG_SCHEDULE")) {
    if (std::string(e) == "fused") return GSG_SCHEDULE_FUSED;
    if (std::string(e) == "streams") return GSG_SCHEDULE_STREAMS;
  }
  std::call_once(g_tune_loaded, load_tune_file);
  std::lock_guard<std::mutex> g(g_tune_mu);
  return g_tune_schedule == GSG_SCHEDULE_FUSED ? GSG_SCHEDULE_FUSED : GSG_SCHEDULE_STREAMS;
}

// the segmented kernel's alternative entries per round exists for the float operator of
// the segmented shapes only
static bool tile_alt_ok(int D, int k, bool fp32) { return fp32 && seg_ok(D, k) && seg_cpw(D, k) > 1; }

// The variant a grid runs: GSG_KERNEL_PLAN, GSG_KERNEL_TILE or GSG_KERNEL_TILE_ALT
// (GSG_KERNEL=plan / tile in the environment, else the tuning table, else the built-in
// rule: the tile variant wherever it applies).
static int grid_variant(const Tables& T, bool fp32) {
  if (!T.tile_ok) return GSG_KERNEL_PLAN;
  if (const char* kv = std::getenv("GSG_KERNEL")) {
    if (std::string(kv) == "plan") return GSG_KERNEL_PLAN;
    if (std::string(kv) == "tile") return GSG_KERNEL_TILE;
  }
  std::call_once(g_tune_loaded, load_tune_file);
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tune.find(std::make_tuple(T.D, T.k, T.n, fp32 ? 1 : 0));
  if (it == g_tune.end() || it->second == GSG_KERNEL_AUTO) return GSG_KERNEL_TILE;
  if (it->second == GSG_KERNEL_TILE_ALT && !tile_alt_ok(T.D, T.k, fp32)) return GSG_KERNEL_TILE;
  return it->second;
}
static bool use_tile(const Tables& T, bool fp32) { return grid_variant(T, fp32) != GSG_KERNEL_PLAN; }

// ---------------------------------------------------------------------------------
// one solve's device state

struct SolveState {
  std::shared_ptr<Tables> T;
  int kind = GSG_OP_TRANSPORT;
  bool fp32 = false;
  int64_t N = 0, len = 0;
  DevMem y, acc, xs0, xs1, xs2, w;
  SolveDesc desc{};

  // entries of a stage-input buffer: the state length, plus u's rounding remainder in
  // the FP32 wave form (the gradient pass reads u as hi + lo floats)
  static int64_t xs_len(int kind, int64_t N, bool fp32) {
    const int64_t len = kind == GSG_OP_WAVE ? 2 * N : N;
    return len + (fp32 && kind == GSG_OP_WAVE ? N : 0);
  }
  static size_t footprint(int kind, int D, int64_t N, bool fp32) {
    const int64_t len = kind == GSG_OP_WAVE ? 2 * N : N;
    const size_t es = fp32 ? 4 : 8;
    const size_t xl = size_t(xs_len(kind, N, fp32));
    size_t b = size_t(len) * 8 * 2 + xl * es * 2;
    if (fp32) b += xl * es;
    if (kind == GSG_OP_WAVE) b += size_t(D) * N * es;
    return b;
  }
  // the stage-1 input of an FP32 solve from its state y (with u's remainder for the wave)
  void to_stage0(hipStream_t s) {
    if (fp32) HIPC(launch_to_f32(y.as<double>(), xs0.as<float>(), len, s, kind == GSG_OP_WAVE ? N : 0));
  }
  size_t bytes() const { return y.bytes + acc.bytes + xs0.bytes + xs1.bytes + xs2.bytes + w.bytes; }

  void init(std::shared_ptr<Tables> tab, const gsg_op_desc& d, bool f32) {
    T = std::move(tab);
    kind = d.kind;
    fp32 = f32;
    N = T->lt.N;
    len = kind == GSG_OP_WAVE ? 2 * N : N;
    const size_t es = fp32 ? 4 : 8;
    const size_t xl = size_t(xs_len(kind, N, fp32));
    y.alloc(len * 8);
    acc.alloc(len * 8);
    if (fp32) xs0.alloc(xl * es);
    xs1.alloc(xl * es);
    xs2.alloc(xl * es);
    if (kind == GSG_OP_WAVE) w.alloc(size_t(T->D) * N * es);
    std::memset(&desc, 0, sizeof(desc));
    desc.t = T->devp[fp32 ? 1 : 0];
    desc.D = T->D; desc.K = T->k; desc.n = T->n; desc.full = T->full;
    desc.N = N;
    desc.sys = kind == GSG_OP_WAVE ? SYS_WAVE : SYS_SUM;
    desc.ndir = T->D;
    desc.dir0 = 0;
    for (int e = 0; e < T->D; ++e) desc.coef[e] = kind == GSG_OP_TRANSPORT ? -d.coef[e] : 1.0;
    desc.y = y.as<double>();
    desc.acc = acc.as<double>();
    desc.xs[0] = xs0.p;
    desc.xs[1] = xs1.p;
    desc.xs[2] = xs2.p;
    desc.w = w.p;
    desc.active = 1;
  }
};

static void check_desc(const gsg_op_desc* d) {
  require(d != nullptr, "desc is NULL");
  require(d->D >= 1 && d->D <= MAXD, "D must be in 1..8");
  require(d->k >= 1 && d->k <= KMAX, "k must be in 1..10");
  require(d->n >= 0 && d->n <= 20, "n must be in 0..20");
  require(d->scheme == GSG_SCHEME_SPARSE || d->scheme == GSG_SCHEME_FULL, "unknown scheme");
  require(d->precision == GSG_PREC_FP64 || d->precision == GSG_PREC_FP32, "unknown precision");
  require(d->kind >= GSG_OP_DERIV && d->kind <= GSG_OP_TRANSPORT, "unknown operator kind");
  if (d->kind == GSG_OP_DERIV) require(d->dir >= 1 && d->dir <= d->D, "dir must be in 1..D");
  for (int e = 0; e < d->D; ++e) require(std::isfinite(d->coef[e]), "coefficients must be finite");
}

// RK4 stage launches of a set of solves of one (D, K) class.
// How a grid's launches run: the kernel and, for the LDS fiber kernel, its LDS layout.
struct LdsShape {
  int32_t kernel = KERNEL_PLAN;
  int32_t bytes = 0, xelems = 0, ments = 0, mslots = 0;
};
// LDS layout of the fiber kernel for batches of at most `slots` input cells and `ents`
// entries: the input region also holds the per-direction results and keeps the
// coefficient region 16-byte aligned.
static LdsShape lds_layout(int D, int K, bool fp32, int32_t slots, int32_t ents) {
  LdsShape L;
  L.kernel = KERNEL_FIBER;
  const int32_t kD = int32_t(Tables::ipow_i(K, D));
  const int32_t cpt = fiber_cpt(D, K);
  // input slots of fiber_slot(k^D) doubles (16-byte DMAs); the results reuse the area
  L.xelems = std::max(slots * fiber_slot(kD), cpt * D * kD);
  L.xelems = (L.xelems + 3) / 4 * 4;
  L.ments = ents;
  L.mslots = slots;
  const int32_t belems = fiber_uniform(D, K) ? 0 : ents * ((K * K) % 2 ? K * K + 1 : K * K);
  L.bytes = (L.xelems + belems) * (fp32 ? 4 : 8) + 4 * (slots + ents + cpt * D + 1);
  return L;
}
// The launch shape of a grid: the plan kernel unless the tuned choice is the tile
// variant, which is the segmented kernel (FP32 seg_ok shapes, FP64 k = 2), the direct
// kernel where it applies, and the LDS fiber kernel otherwise.
static LdsShape lds_shape(const Tables& T, bool fp32) {
  const int var = grid_variant(T, fp32);
  if (var == GSG_KERNEL_PLAN) return LdsShape{};
  if (fp32 && seg_ok(T.D, T.k) && seg_cpw(T.D, T.k) > 1) {  // one cell a wavefront: the direct kernel
    LdsShape L;
    L.kernel = var == GSG_KERNEL_TILE_ALT ? KERNEL_SEG_ALT : KERNEL_SEG;
    return L;
  }
  if (!fp32 && seg64_ok(T.D, T.k)) {  // FP64, k = 2: several cells a wavefront (k_seg64)
    LdsShape L;
    L.kernel = KERNEL_SEG;
    return L;
  }
  if (fp32 ? direct_ok(T.D, T.k) : direct64_ok(T.D, T.k)) {
    LdsShape L;
    L.kernel = KERNEL_DIRECT;
    return L;
  }
  if (fp32) return LdsShape{};  // the LDS fiber kernel is FP64 only: the plan kernel
  return lds_layout(T.D, T.k, fp32, T.tplan[fp32].max_slots, T.tplan[fp32].max_ents);
}
// The FP64 segmented kernel walks its cells' lists in lockstep rounds, one entry a round:
// it needs a launch of many cells to hide each round's latency.  Measured at D = 4, k = 2,
// n = 8: 1.7x the plan kernel on 190K cells of copies, 0.8x on one solve's 10.5K.  A launch
// of fewer cells runs the plan kernel.
constexpr int64_t SEG64_MIN_CELLS = 65536;
static LdsShape lds_shape(const Tables& T, bool fp32, int64_t launch_cells) {
  const LdsShape L = lds_shape(T, fp32);
  if (!fp32 && L.kernel == KERNEL_SEG && launch_cells < SEG64_MIN_CELLS) return LdsShape{};
  return L;
}
// tiles of a launch shape: tile-plan tiles for the fiber kernel, TILE-DoF tiles otherwise
static const std::vector<Tile>& tiles_of(const Tables& T, const LdsShape& L) {
  const bool single = L.kernel == KERNEL_DIRECT || L.kernel == KERNEL_SEG || L.kernel == KERNEL_SEG_ALT;
  return L.kernel == KERNEL_FIBER ? T.ttiles : single ? T.dtiles : T.tiles;
}
// tiles per workgroup of a launch shape
static size_t tiles_per_wg(const LdsShape& L, int D, int K) {
  if (L.kernel == KERNEL_DIRECT) return DIRECT_CPW;
  if (L.kernel == KERNEL_SEG || L.kernel == KERNEL_SEG_ALT) return size_t(SEG_WAVES * seg_cpw(D, K));
  return 1;
}
// tiles a wavefront takes together (they must belong to one solve)
static size_t tiles_per_wave(const LdsShape& L, int D, int K) {
  return (L.kernel == KERNEL_SEG || L.kernel == KERNEL_SEG_ALT) ? size_t(seg_cpw(D, K)) : 1;
}
static void set_lds(LaunchArgs& a, const LdsShape& L) {
  a.kernel = L.kernel;
  a.lds_bytes = L.bytes;
  a.lds_xelems = L.xelems;
  a.lds_ments = L.ments;
  a.lds_mslots = L.mslots;
}

static void launch_rk_step(int D, int K, bool fp32, int sys, const SolveDesc* dsolves, const Tile* dtiles,
                           int ntiles, const LdsShape& lds, hipStream_t s) {
  for (int stage = 1; stage <= 4; ++stage) {
    LaunchArgs a{};
    a.solves = dsolves;
    a.tiles = dtiles;
    a.ntiles = ntiles;
    a.stage = stage;
    set_lds(a, lds);
    if (sys == SYS_WAVE) {
      a.op = OP_RK_WAVE_GRAD;
      HIPC(launch_op(D, K, fp32, a, s));
      a.op = OP_RK_WAVE_SUM;
      HIPC(launch_op(D, K, fp32, a, s));
    } else {
      a.op = OP_RK_SUM;
      HIPC(launch_op(D, K, fp32, a, s));
    }
  }
}

}  // namespace gsg

using namespace gsg;

// ---------------------------------------------------------------------------------
// operator handle

// GSG_GRAPH=0 launches every kernel eagerly instead of replaying captured RK4 steps
static bool graphs_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("GSG_GRAPH");
    return !(e && std::string(e) == "0");
  }();
  return on;
}

// Captured RK4 steps: a graph of CHUNK steps and one of a single step, replayed to cover
// any step count (CHUNK steps per graph launch amortise the launch cost of small
// grids).  With `lazy`, the first call after a reset runs eagerly and captures nothing,
// so a launch set that is used once (a short-lived batch composition) pays no capture;
// otherwise a call of at least 2 CHUNK steps captures at once.  The graphs hold kernel
// and event nodes only (no memset or copy nodes).
struct StepGraphs {
  static constexpr int CHUNK = 8;
  hipGraph_t g[2] = {nullptr, nullptr};
  hipGraphExec_t e[2] = {nullptr, nullptr};  // [0]: 1 step, [1]: CHUNK steps
  bool used = false;
  void reset() {
    CaptureGuard cg(g_capture_mu);
    for (int i = 0; i < 2; ++i) {
      if (e[i]) (void)hipGraphExecDestroy(e[i]);
      if (g[i]) (void)hipGraphDestroy(g[i]);
      e[i] = nullptr;
      g[i] = nullptr;
    }
    used = false;
  }
  ~StepGraphs() { reset(); }
  // run `n` steps on stream s; enqueue(k, stream) issues k steps eagerly
  template <typename F>
  void run(int64_t n, hipStream_t s, bool lazy, F&& enqueue) {
    if (n <= 0) return;
    const bool first = !used;
    used = true;
    if (!graphs_enabled() || (first && (lazy || n < 2 * CHUNK))) {
      enqueue(n, s);
      return;
    }
    for (int i = 0; i < 2; ++i) {
      if (e[i]) continue;
      CaptureGuard cg(g_capture_mu);
      HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      try {
        enqueue(i ? CHUNK : 1, s);
      } catch (...) {
        // end the broken capture so the stream is usable again, then report
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(s, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        throw;
      }
      HIPC(hipStreamEndCapture(s, &g[i]));
      HIPC(hipGraphInstantiate(&e[i], g[i], nullptr, nullptr, 0));
      // move the executable graph to the device now, so its first launch costs what the
      // later ones do
      HIPC(hipGraphUpload(e[i], s));
    }
    for (; n >= CHUNK; n -= CHUNK) HIPC(hipGraphLaunch(e[1], s));
    for (; n > 0; --n) HIPC(hipGraphLaunch(e[0], s));
  }
};

// A sparse-grid basis transform on the device: the 1-D transform's block table applied
// along direction 0, then 1, ..., each sweep one k_xsweep launch over the grid (and over
// several vectors at once).  FP64.  Per direction, cells whose entry list is long (the
// coarse output cells of m2n / n2m couple to every finer nested cell) get a workgroup
// each and go first; the other cells are cut into TILE-DoF ranges.
struct XForm {
  // entries above which a cell is split over a workgroup; measured on the 2 + 2
  // dimensional Vlasov right-hand side (k = 2, n = 8): 8 beats 32 by 6%, 4 loses 14%
  static constexpr int32_t LONG = 8;
  static_assert(LONG == XELL, "the fixed-width table holds every short list");
  std::shared_ptr<Tables> T;
  DevMem descs;  // D SolveDescs, sweep d uses descs[d]
  DevMem tiles;  // the D directions' tile lists, one after another
  // k = 2 with long lists split over workgroups (so every short list has at most LONG
  // entries): per direction, XELL (input cell offset, block) pairs a cell, the cell's list
  // and then block -1 (long-list cells: all -1), so a lane reads its entries without
  // waiting on the list bounds
  DevMem ell;
  int64_t ncell = 0;
  std::vector<int32_t> toff, tcnt;
  void init(int D, int k, int n, int full, int tkind) {
    T = get_tables(D, k, n, full, tkind);
    std::vector<SolveDesc> ds(static_cast<size_t>(D));
    for (int d = 0; d < D; ++d) {
      SolveDesc& a = ds[size_t(d)];
      std::memset(&a, 0, sizeof(a));
      a.t = T->devp[0];
      a.D = D; a.K = k; a.n = n; a.full = full;
      a.N = T->lt.N;
      a.sys = SYS_SUM;
      a.active = 1;
      a.ndir = 1;
      a.dir0 = d;
      a.coef[d] = 1.0;
    }
    upload(descs, ds);
    const int64_t KD = Tables::ipow_i(k, D);
    ncell = T->lt.N / KD;
    const std::vector<int32_t>& ptr = T->plan.ptr;
    const int32_t maxparts = KD <= TILE / 2 ? int32_t(TILE / KD) : 1;
    if (k == 2 && maxparts > 1) {
      const std::vector<int32_t>& ent = T->plan.ent;
      std::vector<int32_t> e(size_t(D) * size_t(ncell) * 2 * XELL);
      for (int d = 0; d < D; ++d)
        for (int64_t g = 0; g < ncell; ++g) {
          const int32_t p0 = ptr[size_t(g * D + d)];
          const int32_t len = ptr[size_t(g * D + d + 1)] - p0;
          int32_t* row = e.data() + (size_t(d) * size_t(ncell) + size_t(g)) * 2 * XELL;
          for (int u = 0; u < XELL; ++u) {
            const bool in = len <= LONG && u < len;
            row[2 * u] = in ? ent[size_t(2 * (p0 + u))] : 0;
            row[2 * u + 1] = in ? ent[size_t(2 * (p0 + u) + 1)] : -1;
          }
        }
      upload(ell, e);
    }
    std::vector<Tile> all;
    for (int d = 0; d < D; ++d) {
      std::vector<Tile> heavy, light;
      int64_t run = -1;  // first DoF of the current run of short-list cells
      auto flush = [&](int64_t stop) {
        for (int64_t a0 = run; run >= 0 && a0 < stop; a0 += TILE)
          light.push_back(Tile{0, 0, a0, std::min<int64_t>(a0 + TILE, stop), 1});
        run = -1;
      };
      for (int64_t g = 0; g < ncell; ++g) {
        const int32_t len = ptr[size_t(g * D + d + 1)] - ptr[size_t(g * D + d)];
        if (maxparts > 1 && len > LONG) {
          flush(g * KD);
          const int32_t parts = std::min<int32_t>(maxparts, (len + 3) / 4);  // about 4 entries a lane
          heavy.push_back(Tile{0, 0, g * KD, (g + 1) * KD, parts});
        } else if (run < 0) {
          run = g * KD;
        }
      }
      flush(ncell * KD);
      while (heavy.size() % NXCD) heavy.push_back(Tile{0, 0, 0, 0, 2});  // keeps the XCD runs aligned
      const std::vector<Tile> pl = xcd_split(light);
      toff.push_back(int32_t(all.size()));
      tcnt.push_back(int32_t(heavy.size() + pl.size()));
      all.insert(all.end(), heavy.begin(), heavy.end());
      all.insert(all.end(), pl.begin(), pl.end());
    }
    upload(tiles, all);
  }
  // y = T x for nvec vectors of N elements stored one after another, with two scratch
  // areas of as many doubles vectors (all distinct from x and y).  x32: the inputs are
  // float (read by the first sweep).  comb: the last sweep writes the Vlasov combination
  // of the nvec results into y (one vector).
  void apply(const void* x, double* y, double* ta, double* tb, hipStream_t s, int nvec = 1, bool x32 = false,
             const XComb* comb = nullptr, const XRk* rk = nullptr) const {
    const int D = T->D;
    const void* cur = x;
    for (int d = 0; d < D; ++d) {
      double* next = d == D - 1 ? y : (d % 2 ? tb : ta);
      LaunchArgs a{};
      a.solves = descs.as<SolveDesc>() + d;
      a.tiles = tiles.as<Tile>() + toff[size_t(d)];
      a.ntiles = tcnt[size_t(d)];
      a.xin = cur;
      a.yout = next;
      const int32_t* el = ell.p ? ell.as<int32_t>() + size_t(d) * size_t(ncell) * 2 * XELL : nullptr;
      HIPC(launch_xsweep(T->k, nvec, a, s, x32 && d == 0, d == D - 1 ? comb : nullptr, d == D - 1 ? rk : nullptr,
                         el));
      cur = next;
    }
  }
};

struct gsg_op {
  gsg_op_desc d{};
  bool fp32 = false;
  bool is_csr = false;
  int device = 0;
  // every device launch of the handle runs on its own stream; ev_in / ev_out order that
  // stream after the caller's stream and the caller's stream after it
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  std::shared_ptr<Tables> T;
  int64_t N = 0, len = 0;
  // apply path
  SolveDesc adesc{};
  DevMem d_adesc, d_tiles;  // tiles of the launch shape, XCD-split
  int32_t n_tiles = 0;
  LdsShape shape;  // kernel and LDS layout of the apply and RK launches
  DevMem xbuf, ybuf, x32, wbuf;
  // rk path
  std::unique_ptr<SolveState> rk;
  DevMem d_rkdesc;
  double rk_dt = 0.0;      // dt in d_rkdesc
  bool rk_desc_set = false;
  StepGraphs rkgraphs;  // the launches read dt and the state through the descriptor, so
                        // captured steps stay valid for every later call
  // csr
  DevMem rp, ci, v64, v32;
  DevCsr csr{};
  // basis transform (kind GSG_OP_TRANSFORM) / the Vlasov right-hand side (GSG_OP_VLASOV)
  XForm xf;
  DevMem ta, tb;  // sweep scratch
  int32_t Dx = 0;
  bool has_force = false;
  XForm m2n, n2p, p2n, n2m;
  DevMem vpt, fpt, cbuf, t1;
  DevMem vy, vacc, vx, vr;  // Vlasov RK4 state, stage input, right-hand side
  double vdt = 0.0;         // dt the captured Vlasov steps were recorded with
  StepGraphs vgraphs;
};

// the Vlasov right-hand side y = F(x) (x, y: N doubles on the device); x32_ready: the
// FP32 form's float copy of x is already in op->x32; rk (stage > 0): the last sweep
// applies that RK4 stage update instead of writing y
static void vlasov_apply(gsg_op* op, const double* x, double* y, hipStream_t s, bool x32_ready = false,
                         const XRk* rk = nullptr) {
  const Tables& T = *op->T;
  const int D = T.D, Dx = op->Dx;
  const int64_t N = op->N;
  const void* xin = x;
  if (op->fp32) {
    if (!x32_ready) HIPC(launch_to_f32(x, op->x32.as<float>(), N, s));
    xin = op->x32.p;
  }
  LaunchArgs a{};
  a.solves = op->d_adesc.as<SolveDesc>();
  a.tiles = op->d_tiles.as<Tile>();
  a.ntiles = op->n_tiles;
  set_lds(a, op->shape);
  a.op = OP_APPLY_GRAD;  // all 2 Dx directional derivatives, D_d f
  a.xin = xin;
  a.yout = op->wbuf.p;
  HIPC(launch_op(D, T.k, op->fp32, a, s));
  double* ta = op->ta.as<double>();
  double* tb = op->tb.as<double>();
  double* t1 = op->t1.as<double>();
  // point values of the Dx spatial derivatives (and of the Dx velocity derivatives with
  // a force): the derivative vectors w[0..nv) together through m2n (reading the FP32
  // derivative directly), then n2p, whose last sweep forms -sum v_d P_d + sum F_d Q_d
  const int nv = op->has_force ? D : Dx;
  const XComb comb{op->vpt.as<double>(), op->has_force ? op->fpt.as<double>() : nullptr, Dx, nv};
  op->m2n.apply(op->wbuf.p, t1, ta, tb, s, nv, op->fp32);
  op->n2p.apply(t1, op->cbuf.as<double>(), ta, tb, s, nv, false, &comb);
  op->p2n.apply(op->cbuf.as<double>(), t1, ta, tb, s);
  op->n2m.apply(t1, y, ta, tb, s, 1, false, nullptr, rk);
}

// n RK4 steps of the Vlasov system on op->vy (dt fixed: the captured steps hold it)
static void vlasov_rk_steps(gsg_op* op, double dt, int64_t n, hipStream_t s) {
  if (dt != op->vdt) {
    op->vgraphs.reset();
    op->vdt = dt;
  }
  // FP32: the stage update also writes the float copy of the next stage input (and of
  // the new state), so only the first stage of a run converts
  float* x32 = op->fp32 ? op->x32.as<float>() : nullptr;
  op->vgraphs.run(n, s, false, [&](int64_t k, hipStream_t st) {
    for (int64_t i = 0; i < k; ++i)
      for (int stage = 1; stage <= 4; ++stage) {
        // the last n2m sweep applies the stage update (no right-hand-side vector)
        const XRk rk{stage, dt, op->vy.as<double>(), op->vacc.as<double>(), op->vx.as<double>(), x32};
        vlasov_apply(op, stage == 1 ? op->vy.as<double>() : op->vx.as<double>(), op->vr.as<double>(), st,
                     i > 0 || stage > 1, &rk);
      }
  });
}

// Runs fn(op->stream) ordered after the work already enqueued on the caller's stream cs
// (NULL: the null stream) and before the work the caller enqueues on cs afterwards.  The
// handle's scratch buffers and captured steps are only touched on op->stream, so calls
// made on different caller streams are also ordered among themselves.
template <typename F>
static void on_op_stream(gsg_op* op, void* cs, F&& fn) {
  hipStream_t c = static_cast<hipStream_t>(cs);
  HIPC(hipEventRecord(op->ev_in, c));
  HIPC(hipStreamWaitEvent(op->stream, op->ev_in, 0));
  fn(op->stream);
  HIPC(hipEventRecord(op->ev_out, op->stream));
  HIPC(hipStreamWaitEvent(c, op->ev_out, 0));
}

static void op_common_init(gsg_op* op) {
  HIPC(hipGetDevice(&op->device));
  CaptureGuard cg(g_capture_mu);
  HIPC(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking));
  HIPC(hipEventCreateWithFlags(&op->ev_in, hipEventDisableTiming));
  HIPC(hipEventCreateWithFlags(&op->ev_out, hipEventDisableTiming));
}

static void op_apply_dev(gsg_op* op, const double* x, double* y, hipStream_t s) {
  if (op->d.kind == GSG_OP_TRANSFORM) {
    op->xf.apply(x, y, op->ta.as<double>(), op->tb.as<double>(), s);
    return;
  }
  if (op->d.kind == GSG_OP_VLASOV) {
    vlasov_apply(op, x, y, s);
    return;
  }
  if (op->is_csr) {
    const void* xin = x;
    if (op->fp32) {
      HIPC(launch_to_f32(x, op->x32.as<float>(), op->csr.n, s));
      xin = op->x32.p;
    }
    HIPC(launch_csr(op->csr, op->fp32, xin, y, s));
    return;
  }
  const Tables& T = *op->T;
  const int D = T.D, K = T.k;
  const void* xin = x;
  // the float gradient pass of the Laplacian and the wave system also reads u's rounding
  // remainder, stored after the input (at op->len)
  const bool two_pass = op->d.kind == GSG_OP_LAPLACIAN || op->d.kind == GSG_OP_WAVE;
  if (op->fp32) {
    HIPC(launch_to_f32(x, op->x32.as<float>(), op->len, s, two_pass ? op->N : 0));
    xin = op->x32.p;
  }
  LaunchArgs a{};
  a.solves = op->d_adesc.as<SolveDesc>();
  a.tiles = op->d_tiles.as<Tile>();
  a.ntiles = op->n_tiles;
  set_lds(a, op->shape);
  switch (op->d.kind) {
    case GSG_OP_DERIV:
    case GSG_OP_GRAD_SUM:
    case GSG_OP_TRANSPORT:
      a.op = OP_APPLY_SUM;
      a.xin = xin;
      a.yout = y;
      a.instride = 0;
      HIPC(launch_op(D, K, op->fp32, a, s));
      break;
    case GSG_OP_LAPLACIAN:
      a.op = OP_APPLY_GRAD;
      a.xin = xin;
      a.yout = op->wbuf.p;
      a.split = op->fp32 ? op->len : 0;
      HIPC(launch_op(D, K, op->fp32, a, s));
      a.split = 0;
      a.op = OP_APPLY_SUM;
      a.xin = op->wbuf.p;
      a.yout = y;
      a.instride = op->N;
      HIPC(launch_op(D, K, op->fp32, a, s));
      break;
    case GSG_OP_WAVE:
      a.op = OP_APPLY_GRAD;
      a.xin = xin;
      a.yout = op->wbuf.p;
      a.split = op->fp32 ? op->len : 0;
      HIPC(launch_op(D, K, op->fp32, a, s));
      a.split = 0;
      a.op = OP_APPLY_WAVE_SUM;
      a.xin = xin;
      a.xin2 = op->wbuf.p;
      a.yout = y;
      HIPC(launch_op(D, K, op->fp32, a, s));
      break;
    default:
      throw Error(GSG_EINVAL, "operator kind cannot be applied");
  }
}

extern "C" {

const char* gsg_last_error(void) { return g_last_error.c_str(); }

const char* gsg_version(void) { return "gsg-mi355x 0.1 (gfx950)"; }

int gsg_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int gsg_tune_set(int32_t D, int32_t k, int32_t n, int32_t precision, int32_t variant) {
  return guarded([&] {
    require(D >= 1 && D <= MAXD && k >= 1 && k <= KMAX && n >= 0 && n <= 20, "invalid (D, k, n)");
    require(precision == GSG_PREC_FP64 || precision == GSG_PREC_FP32, "unknown precision");
    require(variant == GSG_KERNEL_AUTO || variant == GSG_KERNEL_PLAN || variant == GSG_KERNEL_TILE ||
                variant == GSG_KERNEL_TILE_ALT,
            "unknown kernel variant");
    std::call_once(g_tune_loaded, load_tune_file);
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[std::make_tuple(D, k, n, precision)] = variant;
  });
}

int gsg_tune_get(int32_t D, int32_t k, int32_t n, int32_t precision) {
  int v = GSG_KERNEL_PLAN;
  const int rc = guarded([&] {
    require(D >= 1 && D <= MAXD && k >= 1 && k <= KMAX && n >= 0 && n <= 20, "invalid (D, k, n)");
    require(precision == GSG_PREC_FP64 || precision == GSG_PREC_FP32, "unknown precision");
    Tables probe;
    probe.D = D;
    probe.k = k;
    probe.n = n;
    probe.tile_ok = fiber_ok(D, k);
    v = grid_variant(probe, precision == GSG_PREC_FP32);
  });
  return rc < 0 ? rc : v;
}

int64_t gsg_get_size(int32_t D, int32_t k, int32_t n, int32_t scheme) {
  if (D < 1 || D > 16 || k < 1 || k > KMAX || n < 0 || n > 30 || (scheme != 0 && scheme != 1)) {
    g_last_error = "invalid (D, k, n, scheme)";
    return GSG_EINVAL;
  }
  return layout_size(D, k, n, scheme == GSG_SCHEME_FULL);
}

int gsg_v2dref(int32_t D, int32_t k, int32_t n, int32_t scheme, int32_t* levels, int32_t* cells, int32_t* modes,
               int64_t len) {
  return guarded([&] {
    require(levels && cells && modes, "NULL output");
    Layout lt = build_layout(D, k, n, scheme == GSG_SCHEME_FULL);
    require(len == lt.N, "len must equal get_size");
    layout_v2dref(lt, levels, cells, modes);
  });
}

int gsg_dg_basis(int32_t k, double* out) {
  return guarded([&] {
    require(out != nullptr, "NULL output");
    auto b = dg_basis(k);
    std::copy(b.begin(), b.end(), out);
  });
}

int gsg_dlf_1d(int32_t k, int32_t n, double* out, int64_t len) {
  return guarded([&] {
    require(out != nullptr, "NULL output");
    require(n >= 0 && n <= 12, "n must be in 0..12 for the dense export");
    const int64_t N1 = int64_t(k) << n;
    require(len == N1 * N1, "len must be (k 2^n)^2");
    Op1D op = build_op1d(k, n);
    auto Dm = op1d_dense(op);
    std::copy(Dm.begin(), Dm.end(), out);
  });
}

int gsg_d_matrix_csc(int32_t D, int32_t d, int32_t k, int32_t n, int32_t scheme, int64_t* colptr, int64_t* rowval,
                     double* nzval, int64_t* nnz) {
  return guarded([&] {
    require(nnz != nullptr, "nnz is NULL");
    require(d >= 1 && d <= D, "d must be in 1..D");
    Layout lt = build_layout(D, k, n, scheme == GSG_SCHEME_FULL);
    Op1D op = build_op1d(k, n);
    Csc A = d_matrix_csc(lt, op, d - 1);
    if (!nzval) {
      *nnz = int64_t(A.nzval.size());
      return;
    }
    require(*nnz == int64_t(A.nzval.size()), "nnz does not match the matrix");
    require(colptr && rowval, "NULL output");
    for (size_t t = 0; t < A.colptr.size(); ++t) colptr[t] = A.colptr[t] + 1;
    for (size_t t = 0; t < A.rowval.size(); ++t) rowval[t] = A.rowval[t] + 1;
    std::copy(A.nzval.begin(), A.nzval.end(), nzval);
  });
}

int gsg_op_create(const gsg_op_desc* desc, gsg_op** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    check_desc(desc);
    auto op = std::make_unique<gsg_op>();
    op->d = *desc;
    op->fp32 = desc->precision == GSG_PREC_FP32;
    op->T = get_tables(desc->D, desc->k, desc->n, desc->scheme);
    op_common_init(op.get());
    const Tables& T = *op->T;
    op->N = T.lt.N;
    op->len = desc->kind == GSG_OP_WAVE ? 2 * op->N : op->N;
    SolveDesc& a = op->adesc;
    std::memset(&a, 0, sizeof(a));
    a.t = T.devp[op->fp32 ? 1 : 0];
    a.D = T.D; a.K = T.k; a.n = T.n; a.full = T.full;
    a.N = op->N;
    a.sys = desc->kind == GSG_OP_WAVE ? SYS_WAVE : SYS_SUM;
    a.active = 1;
    a.ndir = T.D;
    a.dir0 = 0;
    for (int e = 0; e < T.D; ++e) {
      switch (desc->kind) {
        case GSG_OP_DERIV: a.coef[e] = 1.0; break;
        case GSG_OP_GRAD_SUM: a.coef[e] = desc->coef[e]; break;
        case GSG_OP_TRANSPORT: a.coef[e] = -desc->coef[e]; break;
        default: a.coef[e] = 1.0; break;
      }
    }
    if (desc->kind == GSG_OP_DERIV) {
      a.ndir = 1;
      a.dir0 = desc->dir - 1;
    }
    op->d_adesc.alloc(sizeof(SolveDesc));
    {
      CaptureGuard cg(g_capture_mu);
      HIPC(hipMemcpy(op->d_adesc.p, &a, sizeof(SolveDesc), hipMemcpyHostToDevice));
    }
    // the fiber and direct kernels apply every direction: a single derivative runs the plan kernel
    if (desc->kind != GSG_OP_DERIV) op->shape = lds_shape(T, op->fp32, T.lt.N / Tables::ipow_i(T.k, T.D));
    const std::vector<Tile> pt = xcd_split(tiles_of(T, op->shape), tiles_per_wg(op->shape, T.D, T.k),
                                           tiles_per_wave(op->shape, T.D, T.k));
    op->n_tiles = int32_t(pt.size());
    upload(op->d_tiles, pt);
    const size_t es = op->fp32 ? 4 : 8;
    const bool two_pass = desc->kind == GSG_OP_LAPLACIAN || desc->kind == GSG_OP_WAVE;
    if (op->fp32) op->x32.alloc((op->len + (two_pass ? op->N : 0)) * 4);  // two passes: and u's remainder
    if (desc->kind == GSG_OP_LAPLACIAN || desc->kind == GSG_OP_WAVE) op->wbuf.alloc(size_t(T.D) * op->N * es);
    *out = op.release();
  });
}

int gsg_csc_create(int64_t m, int64_t n, const int64_t* colptr, const int64_t* rowval, const double* nzval,
                   int64_t nnz, int32_t precision, gsg_op** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    require(m > 0 && n > 0 && m == n, "only square matrices are supported");
    require(m < (int64_t(1) << 31), "matrix too large");
    require(nnz >= 0 && colptr && (nnz == 0 || (rowval && nzval)), "bad CSC arrays");
    require(precision == GSG_PREC_FP64 || precision == GSG_PREC_FP32, "unknown precision");
    require(colptr[0] == 1 && colptr[n] == nnz + 1, "colptr must be 1-based and end at nnz+1");
    ensure_device();
    // transpose to CSR with ascending columns per row
    std::vector<int64_t> rp(m + 1, 0);
    for (int64_t j = 0; j < n; ++j) {
      require(colptr[j + 1] >= colptr[j], "colptr must be non-decreasing");
      for (int64_t p = colptr[j] - 1; p < colptr[j + 1] - 1; ++p) {
        require(rowval[p] >= 1 && rowval[p] <= m, "rowval out of range");
        ++rp[rowval[p]];
      }
    }
    for (int64_t i = 0; i < m; ++i) rp[i + 1] += rp[i];
    std::vector<int64_t> pos(rp.begin(), rp.end() - 1);
    std::vector<int32_t> ci(std::max<int64_t>(nnz, 1));
    std::vector<double> v(std::max<int64_t>(nnz, 1));
    for (int64_t j = 0; j < n; ++j)
      for (int64_t p = colptr[j] - 1; p < colptr[j + 1] - 1; ++p) {
        const int64_t q = pos[rowval[p] - 1]++;
        ci[q] = int32_t(j);
        v[q] = nzval[p];
      }
    auto op = std::make_unique<gsg_op>();
    op->is_csr = true;
    op->fp32 = precision == GSG_PREC_FP32;
    op->d.kind = GSG_OP_CSR;
    op->d.precision = precision;
    op->N = m;
    op->len = m;
    op_common_init(op.get());
    upload(op->rp, rp);
    upload(op->ci, ci);
    upload(op->v64, v);
    std::vector<float> v32(2 * v.size());  // hi values, then lo values
    for (size_t i = 0; i < v.size(); ++i) {
      v32[i] = float(v[i]);
      v32[v.size() + i] = float(v[i] - double(v32[i]));
    }
    upload(op->v32, v32);
    if (op->fp32) op->x32.alloc(size_t(n) * 4);
    op->csr = DevCsr{m, n, nnz, op->rp.as<int64_t>(), op->ci.as<int32_t>(), op->v64.as<double>(), op->v32.as<float>()};
    *out = op.release();
  });
}

static int tb_kind(int from, int to) {
  if (from == GSG_BASIS_MODAL && to == GSG_BASIS_NODAL) return TAB_M2N;
  if (from == GSG_BASIS_NODAL && to == GSG_BASIS_MODAL) return TAB_N2M;
  if (from == GSG_BASIS_NODAL && to == GSG_BASIS_POINTS) return TAB_N2P;
  if (from == GSG_BASIS_POINTS && to == GSG_BASIS_NODAL) return TAB_P2N;
  throw Error(GSG_EINVAL, "(from, to) must be (MODAL, NODAL), (NODAL, MODAL), (NODAL, POINTS) or (POINTS, NODAL)");
}

int gsg_transform_1d(int32_t k, int32_t n, int32_t from, int32_t to, double* out, int64_t len) {
  return guarded([&] {
    require(out != nullptr, "NULL output");
    require(nodal_k_ok(k), "the nodal basis exists for k in {2, 3, 5}");
    require(n >= 0 && n <= 12, "n must be in 0..12 for the dense export");
    (void)tb_kind(from, to);
    const int64_t N1 = int64_t(k) << n;
    require(len == N1 * N1, "len must be (k 2^n)^2");
    const std::vector<double> M = transform_1d(k, n, from, to);
    std::copy(M.begin(), M.end(), out);
  });
}

int gsg_points_1d(int32_t k, int32_t n, double* x, int64_t len) {
  return guarded([&] {
    require(x != nullptr, "NULL output");
    require(nodal_k_ok(k), "the nodal basis exists for k in {2, 3, 5}");
    require(n >= 0 && n <= 12, "n must be in 0..12");
    require(len == (int64_t(k) << n), "len must be k 2^n");
    const std::vector<double> p = points_1d(k, n);
    std::copy(p.begin(), p.end(), x);
  });
}

int gsg_transform_create(int32_t D, int32_t k, int32_t n, int32_t scheme, int32_t from, int32_t to, gsg_op** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    require(D >= 1 && D <= MAXD, "D must be in 1..8");
    require(nodal_k_ok(k), "the nodal basis exists for k in {2, 3, 5}");
    require(n >= 0 && n <= 12, "n must be in 0..12");
    require(scheme == GSG_SCHEME_SPARSE || scheme == GSG_SCHEME_FULL, "unknown scheme");
    const int tk = tb_kind(from, to);
    auto op = std::make_unique<gsg_op>();
    op->d.D = D; op->d.k = k; op->d.n = n; op->d.scheme = scheme;
    op->d.kind = GSG_OP_TRANSFORM;
    op->d.precision = GSG_PREC_FP64;
    op_common_init(op.get());
    op->xf.init(D, k, n, scheme, tk);
    op->T = op->xf.T;
    op->N = op->len = op->T->lt.N;
    op->ta.alloc(size_t(op->N) * 8);
    op->tb.alloc(size_t(op->N) * 8);
    *out = op.release();
  });
}

int gsg_transform_csc(int32_t D, int32_t k, int32_t n, int32_t scheme, int32_t from, int32_t to, int64_t* colptr,
                      int64_t* rowval, double* nzval, int64_t* nnz) {
  // the count call assembles and keeps the matrix for the fill call with the same arguments
  static thread_local std::tuple<int, int, int, int, int, int> key{-1, 0, 0, 0, 0, 0};
  static thread_local Csc kept;
  return guarded([&] {
    require(nnz != nullptr, "nnz is NULL");
    require(D >= 1 && D <= MAXD, "D must be in 1..8");
    require(nodal_k_ok(k), "the nodal basis exists for k in {2, 3, 5}");
    require(n >= 0 && n <= 12, "n must be in 0..12");
    require(scheme == GSG_SCHEME_SPARSE || scheme == GSG_SCHEME_FULL, "unknown scheme");
    (void)tb_kind(from, to);
    const auto want = std::make_tuple(int(D), int(k), int(n), int(scheme), int(from), int(to));
    if (key != want) {
      kept = Csc();
      key = std::make_tuple(-1, 0, 0, 0, 0, 0);
      const Layout lt = build_layout(D, k, n, scheme == GSG_SCHEME_FULL);
      kept = transform_csc(lt, transform_op1d(k, n, from, to));
      key = want;
    }
    if (!nzval) {
      *nnz = int64_t(kept.nzval.size());
      return;
    }
    require(*nnz == int64_t(kept.nzval.size()), "nnz does not match the matrix");
    require(colptr && rowval, "NULL output");
    for (size_t t = 0; t < kept.colptr.size(); ++t) colptr[t] = kept.colptr[t] + 1;
    for (size_t t = 0; t < kept.rowval.size(); ++t) rowval[t] = kept.rowval[t] + 1;
    std::copy(kept.nzval.begin(), kept.nzval.end(), nzval);
    kept = Csc();
    key = std::make_tuple(-1, 0, 0, 0, 0, 0);
  });
}

int gsg_vlasov_create(const gsg_vlasov_desc* desc, const double* F_point, int64_t F_len, gsg_op** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    require(desc != nullptr, "desc is NULL");
    const int Dx = desc->Dx, k = desc->k, n = desc->n, scheme = desc->scheme;
    require(Dx >= 1 && 2 * Dx <= MAXD, "Dx must be in 1..4 (2 Dx phase-space dimensions)");
    require(nodal_k_ok(k), "the nodal basis exists for k in {2, 3, 5}");
    require(n >= 0 && n <= 12, "n must be in 0..12");
    require(scheme == GSG_SCHEME_SPARSE || scheme == GSG_SCHEME_FULL, "unknown scheme");
    require(desc->precision == GSG_PREC_FP64 || desc->precision == GSG_PREC_FP32, "unknown precision");
    const int D = 2 * Dx;
    auto op = std::make_unique<gsg_op>();
    op->d.D = D; op->d.k = k; op->d.n = n; op->d.scheme = scheme;
    op->d.kind = GSG_OP_VLASOV;
    op->d.precision = desc->precision;
    op->fp32 = desc->precision == GSG_PREC_FP32;
    op->Dx = Dx;
    op_common_init(op.get());
    op->T = get_tables(D, k, n, scheme);
    const Tables& T = *op->T;
    const int64_t N = T.lt.N;
    op->N = op->len = N;
    require(F_point == nullptr || F_len == int64_t(Dx) * N, "F_len must be Dx * get_size(2 Dx, k, n)");
    op->has_force = F_point != nullptr;
    // the 2 Dx directional derivatives: the operator's own kernel choice, every direction
    SolveDesc& a = op->adesc;
    std::memset(&a, 0, sizeof(a));
    a.t = T.devp[op->fp32 ? 1 : 0];
    a.D = D; a.K = k; a.n = n; a.full = scheme;
    a.N = N;
    a.sys = SYS_SUM;
    a.active = 1;
    a.ndir = D;
    for (int e = 0; e < D; ++e) a.coef[e] = 1.0;
    upload(op->d_adesc, std::vector<SolveDesc>{a});
    op->shape = lds_shape(T, op->fp32, T.lt.N / Tables::ipow_i(k, D));
    const std::vector<Tile> pt = xcd_split(tiles_of(T, op->shape), tiles_per_wg(op->shape, D, k),
                                           tiles_per_wave(op->shape, D, k));
    op->n_tiles = int32_t(pt.size());
    upload(op->d_tiles, pt);
    op->m2n.init(D, k, n, scheme, TAB_M2N);
    op->n2p.init(D, k, n, scheme, TAB_N2P);
    op->p2n.init(D, k, n, scheme, TAB_P2N);
    op->n2m.init(D, k, n, scheme, TAB_N2M);
    const size_t vb = size_t(N) * 8;
    op->wbuf.alloc(size_t(D) * N * (op->fp32 ? 4 : 8));
    if (op->fp32) op->x32.alloc(size_t(N) * 4);
    const size_t nv = size_t(op->has_force ? D : Dx);  // vectors transformed together
    op->ta.alloc(nv * vb);
    op->tb.alloc(nv * vb);
    op->t1.alloc(nv * vb);
    op->cbuf.alloc(vb);
    if (op->has_force) {
      op->fpt.alloc(size_t(Dx) * vb);
      CaptureGuard cg(g_capture_mu);
      HIPC(hipMemcpy(op->fpt.p, F_point, size_t(Dx) * vb, hipMemcpyHostToDevice));
    }
    // v_point[d] = n2p m2n v_modal[d] (pdes.jl:167-174): v_modal[d] is the tensor product of
    // get_xi_modal(1, 1, k, n) = (1/sqrt(3)) phi_2 along velocity direction Dx+d with the
    // constant 1 along the others, a single coefficient in the level-0 cell
    op->vpt.alloc(size_t(Dx) * vb);
    {
      std::vector<double> vm(size_t(N), 0.0);
      for (int d = 0; d < Dx; ++d) {
        std::fill(vm.begin(), vm.end(), 0.0);
        vm[size_t(Tables::ipow_i(k, Dx + d))] = 1.0 / std::sqrt(3.0);
        CaptureGuard cg(g_capture_mu);
        HIPC(hipMemcpy(op->cbuf.p, vm.data(), vb, hipMemcpyHostToDevice));
        op->m2n.apply(op->cbuf.as<double>(), op->t1.as<double>(), op->ta.as<double>(), op->tb.as<double>(),
                      op->stream);
        op->n2p.apply(op->t1.as<double>(), op->vpt.as<double>() + int64_t(d) * N, op->ta.as<double>(),
                      op->tb.as<double>(), op->stream);
        HIPC(hipStreamSynchronize(op->stream));
      }
    }
    *out = op.release();
  });
}

int64_t gsg_op_len(const gsg_op* op) { return op ? op->len : GSG_EINVAL; }

int gsg_op_apply_device(gsg_op* op, const double* x, double* y, int64_t len, void* stream) {
  return guarded([&] {
    require(op && x && y, "NULL argument");
    require(len == op->len, "len does not match the operator");
    on_op_stream(op, stream, [&](hipStream_t s) { op_apply_dev(op, x, y, s); });
  });
}

int gsg_op_apply(gsg_op* op, const double* x, double* y, int64_t len) {
  return guarded([&] {
    require(op && x && y, "NULL argument");
    require(len == op->len, "len does not match the operator");
    if (!op->xbuf.p) {
      op->xbuf.alloc(size_t(len) * 8);
      op->ybuf.alloc(size_t(len) * 8);
    }
    HIPC(hipMemcpyAsync(op->xbuf.p, x, size_t(len) * 8, hipMemcpyHostToDevice, op->stream));
    op_apply_dev(op, op->xbuf.as<double>(), op->ybuf.as<double>(), op->stream);
    HIPC(hipMemcpyAsync(y, op->ybuf.p, size_t(len) * 8, hipMemcpyDeviceToHost, op->stream));
    HIPC(hipStreamSynchronize(op->stream));
  });
}

static void vlasov_rk_prepare(gsg_op* op, double dt) {
  require(std::isfinite(dt), "dt must be finite");
  if (!op->vy.p) {
    const size_t vb = size_t(op->N) * 8;
    op->vy.alloc(vb);
    op->vacc.alloc(vb);
    op->vx.alloc(vb);
    op->vr.alloc(vb);
  }
}

static void op_rk_prepare(gsg_op* op, double dt) {
  require(!op->is_csr, "RK4 on an imported CSC matrix: use gsg_op_apply in the caller's integrator");
  require(op->d.kind == GSG_OP_TRANSPORT || op->d.kind == GSG_OP_WAVE || op->d.kind == GSG_OP_GRAD_SUM ||
              op->d.kind == GSG_OP_DERIV,
          "RK4 needs a GSG_OP_TRANSPORT, GSG_OP_WAVE, GSG_OP_GRAD_SUM or GSG_OP_DERIV operator");
  require(std::isfinite(dt), "dt must be finite");
  if (!op->rk) {
    op->rk = std::make_unique<SolveState>();
    gsg_op_desc d = op->d;
    op->rk->init(op->T, d, op->fp32);
    // the RK right-hand side is exactly the operator: reuse its coefficients
    op->rk->desc.coef[0] = 0;
    std::memcpy(op->rk->desc.coef, op->adesc.coef, sizeof(op->adesc.coef));
    op->rk->desc.ndir = op->adesc.ndir;
    op->rk->desc.dir0 = op->adesc.dir0;
    op->d_rkdesc.alloc(sizeof(SolveDesc));
  }
  // the descriptor (with dt) changes only with dt; its write is ordered on op->stream
  // after the steps of earlier calls, which read it
  if (!op->rk_desc_set || dt != op->rk_dt) {
    op->rk->desc.dt = dt;
    HIPC(launch_write_desc(op->d_rkdesc.as<SolveDesc>(), op->rk->desc, op->stream));
    op->rk_dt = dt;
    op->rk_desc_set = true;
  }
}

static void op_rk_steps(gsg_op* op, int64_t nsteps, hipStream_t s) {
  const Tables& T = *op->T;
  const int sys = op->d.kind == GSG_OP_WAVE ? SYS_WAVE : SYS_SUM;
  op->rkgraphs.run(nsteps, s, false, [&](int64_t k, hipStream_t st) {
    for (int64_t i = 0; i < k; ++i)
      launch_rk_step(T.D, T.k, op->fp32, sys, op->d_rkdesc.as<SolveDesc>(), op->d_tiles.as<Tile>(), op->n_tiles,
                     op->shape, st);
  });
}

int gsg_op_rk4(gsg_op* op, double* y, int64_t len, double t0, double dt, int64_t nsteps, int64_t save_every,
               double* snapshots) {
  (void)t0;  // autonomous linear systems: the RHS does not depend on t
  return guarded([&] {
    require(op && y, "NULL argument");
    require(len == op->len, "len does not match the operator");
    require(nsteps >= 0, "nsteps must be >= 0");
    if (op->d.kind == GSG_OP_VLASOV) {
      vlasov_rk_prepare(op, dt);
      hipStream_t s = op->stream;
      HIPC(hipMemcpyAsync(op->vy.p, y, size_t(len) * 8, hipMemcpyHostToDevice, s));
      int64_t done = 0, snap = 0;
      while (done < nsteps) {
        int64_t chunk = nsteps - done;
        if (save_every > 0 && snapshots) chunk = std::min(chunk, save_every - done % save_every);
        vlasov_rk_steps(op, dt, chunk, s);
        done += chunk;
        if (save_every > 0 && snapshots && done % save_every == 0) {
          HIPC(hipMemcpyAsync(snapshots + snap * len, op->vy.p, size_t(len) * 8, hipMemcpyDeviceToHost, s));
          ++snap;
        }
      }
      HIPC(hipMemcpyAsync(y, op->vy.p, size_t(len) * 8, hipMemcpyDeviceToHost, s));
      HIPC(hipStreamSynchronize(s));
      return;
    }
    op_rk_prepare(op, dt);
    SolveState& S = *op->rk;
    hipStream_t s = op->stream;
    HIPC(hipMemcpyAsync(S.y.p, y, size_t(len) * 8, hipMemcpyHostToDevice, s));
    S.to_stage0(s);
    int64_t done = 0, snap = 0;
    while (done < nsteps) {
      int64_t chunk = nsteps - done;
      if (save_every > 0 && snapshots) chunk = std::min(chunk, save_every - done % save_every);
      op_rk_steps(op, chunk, s);
      done += chunk;
      if (save_every > 0 && snapshots && done % save_every == 0) {
        HIPC(hipMemcpyAsync(snapshots + snap * len, S.y.p, size_t(len) * 8, hipMemcpyDeviceToHost, s));
        ++snap;
      }
    }
    HIPC(hipMemcpyAsync(y, S.y.p, size_t(len) * 8, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
  });
}

int gsg_op_rk4_device(gsg_op* op, double* y_dev, int64_t len, double dt, int64_t nsteps, void* stream) {
  return guarded([&] {
    require(op && y_dev, "NULL argument");
    require(len == op->len, "len does not match the operator");
    require(nsteps >= 0, "nsteps must be >= 0");
    if (op->d.kind == GSG_OP_VLASOV) {
      vlasov_rk_prepare(op, dt);
      on_op_stream(op, stream, [&](hipStream_t s) {
        HIPC(hipMemcpyAsync(op->vy.p, y_dev, size_t(len) * 8, hipMemcpyDeviceToDevice, s));
        vlasov_rk_steps(op, dt, nsteps, s);
        HIPC(hipMemcpyAsync(y_dev, op->vy.p, size_t(len) * 8, hipMemcpyDeviceToDevice, s));
      });
      return;
    }
    on_op_stream(op, stream, [&](hipStream_t s) {
      op_rk_prepare(op, dt);
      SolveState& S = *op->rk;
      HIPC(hipMemcpyAsync(S.y.p, y_dev, size_t(len) * 8, hipMemcpyDeviceToDevice, s));
      S.to_stage0(s);
      op_rk_steps(op, nsteps, s);
      HIPC(hipMemcpyAsync(y_dev, S.y.p, size_t(len) * 8, hipMemcpyDeviceToDevice, s));
    });
  });
}

// ---- RK4 without allocation -----------------------------------------------------
// gsg_op_rk4_device_noalloc keeps its RK state (descriptor, y, acc, stage inputs, wave
// gradients) in one file-scope device array per device, held until the process ends,
// instead of the handle's lazily allocated state.  On the host the array's name is only
// a stand-in: kernels, copies and fills use the address hipGetSymbolAddress returns for
// the device that is current.
constexpr size_t NOALLOC_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
constexpr int NOALLOC_MAXDEV = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch[NOALLOC_MAXDEV] = {};

// the device address of `scratch` on device dev (current), looked up once per device
static unsigned char* scratch_of(int dev) {
  require(dev >= 0 && dev < NOALLOC_MAXDEV, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch[dev];
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// bytes of scratch an operator's RK state takes: the descriptor, then the vectors of
// SolveState::init, each on 256 bytes
static size_t noalloc_bytes(const gsg_op* op) {
  const size_t es = op->fp32 ? 4 : 8;
  const size_t xl = size_t(SolveState::xs_len(op->d.kind, op->N, op->fp32));
  size_t b = align256(sizeof(SolveDesc)) + 2 * align256(size_t(op->len) * 8) + 2 * align256(xl * es);
  if (op->fp32) b += align256(xl * es);
  if (op->d.kind == GSG_OP_WAVE) b += align256(size_t(op->T->D) * size_t(op->N) * es);
  return b;
}

int gsg_op_destroy(gsg_op* op) {
  if (!op) return GSG_OK;
  return guarded([&] {
    if (op->stream) {
      (void)hipStreamSynchronize(op->stream);
      {
        CaptureGuard cg(g_capture_mu);
        (void)hipStreamDestroy(op->stream);
        if (op->ev_in) (void)hipEventDestroy(op->ev_in);
        if (op->ev_out) (void)hipEventDestroy(op->ev_out);
      }
    }
    delete op;
  });
}

int gsg_op_rk4_device_noalloc(gsg_op* op, double* y_dev, int64_t len, double dt, int64_t nsteps, void* stream) {
  return guarded([&] {
    require(op && y_dev, "NULL argument");
    require(len == op->len, "len does not match the operator");
    require(nsteps >= 0, "nsteps must be >= 0");
    require(std::isfinite(dt), "dt must be finite");
    require(!op->is_csr && (op->d.kind == GSG_OP_TRANSPORT || op->d.kind == GSG_OP_WAVE ||
                            op->d.kind == GSG_OP_GRAD_SUM || op->d.kind == GSG_OP_DERIV),
            "gsg_op_rk4_device_noalloc needs a GSG_OP_TRANSPORT, GSG_OP_WAVE, GSG_OP_GRAD_SUM or GSG_OP_DERIV operator");
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    require(dev == op->device, "the operator's device must be the current device");
    const size_t need = noalloc_bytes(op);
    if (need > NOALLOC_BYTES)
      throw Error(GSG_EINVAL, "the RK state of this operator takes " + std::to_string(need) +
                                  " bytes; the no-allocation scratch holds " + std::to_string(NOALLOC_BYTES) +
                                  " bytes (use gsg_op_rk4_device)");
    unsigned char* base = scratch_of(dev);
    // layout: descriptor, y, acc, xs0 (FP32 only), xs1, xs2, w (wave only)
    const size_t es = op->fp32 ? 4 : 8;
    size_t off = align256(sizeof(SolveDesc));
    auto take = [&](size_t bytes) {
      unsigned char* p = base + off;
      off += align256(bytes);
      return p;
    };
    const Tables& T = *op->T;
    SolveDesc d;
    std::memset(&d, 0, sizeof(d));
    d.t = T.devp[op->fp32 ? 1 : 0];
    d.D = T.D; d.K = T.k; d.n = T.n; d.full = T.full;
    d.N = op->N;
    d.sys = op->d.kind == GSG_OP_WAVE ? SYS_WAVE : SYS_SUM;
    std::memcpy(d.coef, op->adesc.coef, sizeof(d.coef));  // the RK right-hand side is the operator
    d.ndir = op->adesc.ndir;
    d.dir0 = op->adesc.dir0;
    d.dt = dt;
    const size_t xl = size_t(SolveState::xs_len(op->d.kind, op->N, op->fp32));
    d.y = reinterpret_cast<double*>(take(size_t(len) * 8));
    d.acc = reinterpret_cast<double*>(take(size_t(len) * 8));
    d.xs[0] = op->fp32 ? take(xl * es) : nullptr;
    d.xs[1] = take(xl * es);
    d.xs[2] = take(xl * es);
    d.w = op->d.kind == GSG_OP_WAVE ? take(size_t(T.D) * size_t(op->N) * es) : nullptr;
    d.active = 1;
    SolveDesc* dd = reinterpret_cast<SolveDesc*>(base);
    const int sys = d.sys;
    // eager launches only: no graph, no allocation; every buffer the kernels write is
    // written before it is read within the call (acc by stage 1, xs0 by the conversion,
    // w by each stage's gradient pass), so nothing has to be cleared first
    on_op_stream(op, stream, [&](hipStream_t s) {
      HIPC(launch_write_desc(dd, d, s));
      HIPC(hipMemcpyAsync(d.y, y_dev, size_t(len) * 8, hipMemcpyDeviceToDevice, s));
      if (op->fp32)
        HIPC(launch_to_f32(d.y, static_cast<float*>(d.xs[0]), len, s, op->d.kind == GSG_OP_WAVE ? op->N : 0));
      for (int64_t i = 0; i < nsteps; ++i)
        launch_rk_step(T.D, T.k, op->fp32, sys, dd, op->d_tiles.as<Tile>(), op->n_tiles, op->shape, s);
      HIPC(hipMemcpyAsync(y_dev, d.y, size_t(len) * 8, hipMemcpyDeviceToDevice, s));
    });
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// batches of independent solves

namespace gsg {

struct BatchSolve {
  gsg_op_desc d{};
  std::vector<double> y0;
  double dt = 0;
  int64_t nsteps = 0, done = 0;
  SolveState st;
};

// Algorithmic HBM bytes per component DoF of one RK stage launch: the operator input
// is read once (the gathers' re-reads are cache traffic), the stage epilogue reads and
// writes each RK register it touches once.
static double stage_bytes(int sys, int op, int stage, bool fp32, int D) {
  const double es = fp32 ? 4.0 : 8.0;
  double epi_noinput;  // epilogue traffic of one state entry, operator input excluded
  if (stage == 1) epi_noinput = (fp32 ? 8.0 : 0.0) + 8.0 + es;  // fp64 stage 1 reads y as the input
  else if (stage == 4) epi_noinput = 8.0 + 8.0 + (fp32 ? 4.0 : 0.0);
  else epi_noinput = 8.0 + 16.0 + es;
  if (sys == SYS_SUM) return es + epi_noinput;
  if (op == OP_RK_WAVE_GRAD) return es + D * es;
  return D * es + es + 2.0 * epi_noinput + (stage == 1 && !fp32 ? 16.0 : 0.0);
}

// Device-side grouping of resident solves into (D, K, sys) classes with tile lists.
struct Group {
  int D, K, sys;
  std::vector<Tile> tiles;
  DevMem dtiles;
  int64_t dofs = 0;  // component DoFs covered per launch
  LdsShape lds;      // kernel; for the fiber kernel the LDS layout (maxima over the group)
  bool tiled() const { return lds.kernel != KERNEL_PLAN; }
};

struct Scheduler {
  bool fp32 = false;
  hipStream_t stream = nullptr;
  std::vector<SolveDesc> descs;
  DevMem ddescs;
  std::vector<Group> groups;
  // kernel timing
  bool profile = false;
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  struct Pending {
    int D, K, sys;
    bool tiled;
    int64_t dofs;
    double bytes;
    size_t ev;
  };
  std::vector<Pending> pend;
  std::map<std::tuple<int, int, int, bool>, gsg_kernel_stat> stats;
  // Solves of different classes are independent, so each class runs its stages on one
  // of NLANE streams (default priority) and the classes overlap on the GPU: the small
  // classes fill the gaps the large ones leave.  The batch stream forks to the lanes
  // and joins them after every call.  Profiling runs serially on the batch stream so
  // that each kernel's duration is its own.
  static constexpr int NLANE = 8;  // at most; nlanes of them are used (tuned, default 4)
  int nlanes = 4;
  hipStream_t lanes[NLANE] = {};
  hipEvent_t fork = nullptr, joins[NLANE] = {};
  std::vector<int> lane_of;  // per group
  // The fused schedule (FP32): one launch per stage for every class with a fused body
  // (fused[0]: RK_SUM / RK_WAVE_GRAD, then fused[1]: the wave classes' RK_WAVE_SUM),
  // largest classes first; the other classes launch on the batch stream after it.
  int schedule_own = GSG_SCHEDULE_AUTO;  // the batch's own setting
  int schedule = GSG_SCHEDULE_STREAMS;   // what its launches follow
  FusedArgs fused[2] = {};
  DevMem fused_order[2];  // item -> (segment << 24) | workgroup, the classes interleaved
  int fused_pass_a = FUSED_SUM;
  std::vector<size_t> unfused;  // groups outside the fused launches

  ~Scheduler() {
    CaptureGuard cg(g_capture_mu);
    for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
    reset_graph();
    for (int l = 0; l < NLANE; ++l) {
      if (lanes[l]) (void)hipStreamDestroy(lanes[l]);
      if (joins[l]) (void)hipEventDestroy(joins[l]);
    }
    if (fork) (void)hipEventDestroy(fork);
  }

  void init_lanes() {
    if (fork) return;
    for (int l = 0; l < NLANE; ++l) {
      {
        CaptureGuard cg(g_capture_mu);
        HIPC(hipStreamCreateWithFlags(&lanes[l], hipStreamNonBlocking));
      }
      {
        CaptureGuard cg(g_capture_mu);
        HIPC(hipEventCreateWithFlags(&joins[l], hipEventDisableTiming));
      }
    }
    {
      CaptureGuard cg(g_capture_mu);
      HIPC(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    }
  }

  // lanes by estimated cost (DoFs x D x k), largest class first onto the least loaded lane
  void assign_lanes() {
    lane_of.assign(groups.size(), 0);
    std::vector<size_t> order(groups.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    auto cost = [&](size_t i) { return double(groups[i].dofs) * groups[i].D * groups[i].K; };
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost(x) > cost(y); });
    nlanes = tuned_lanes();
    double load[NLANE] = {};
    for (size_t i : order) {
      const int l = int(std::min_element(load, load + nlanes) - load);
      lane_of[i] = l;
      load[l] += cost(i);
    }
  }

  // rebuild descriptors and tiles for the given solves (all marked active): one group
  // per (D, k, system, kernel variant) class.  A class's tiles, solve after solve, are
  // cut into eight contiguous runs of equal length, one per XCD, so every XCD gets the
  // same work and each solve's vectors are touched by one or two XCDs' L2 caches.
  void build(const std::vector<SolveState*>& solves) {
    reset_graph();
    descs.clear();
    groups.clear();
    std::map<std::tuple<int, int, int, int>, size_t> gidx;
    std::vector<std::vector<size_t>> members;
    // cells per (D, k, system) class of the solves whose grid runs the segmented kernel:
    // a class launch of too few cells runs the plan kernel instead (FP64, SEG64_MIN_CELLS)
    std::map<std::tuple<int, int, int>, int64_t> seg_cells;
    for (SolveState* S : solves)
      if (lds_shape(*S->T, fp32).kernel == KERNEL_SEG)
        seg_cells[std::make_tuple(S->T->D, S->T->k, S->desc.sys)] += S->N / Tables::ipow_i(S->T->k, S->T->D);
    for (size_t s = 0; s < solves.size(); ++s) {
      SolveState& S = *solves[s];
      descs.push_back(S.desc);
      const auto ck = std::make_tuple(S.T->D, S.T->k, S.desc.sys);
      const LdsShape L = lds_shape(*S.T, fp32, seg_cells.count(ck) ? seg_cells[ck] : int64_t(1) << 40);
      auto key = std::make_tuple(S.T->D, S.T->k, S.desc.sys, int(L.kernel));
      auto it = gidx.find(key);
      if (it == gidx.end()) {
        gidx[key] = groups.size();
        groups.push_back(Group{S.T->D, S.T->k, S.desc.sys, {}, {}, 0, L});
        members.emplace_back();
        it = gidx.find(key);
      }
      groups[it->second].dofs += S.N;
      members[it->second].push_back(s);
    }
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      Group& g = groups[gi];
      std::vector<Tile> all;
      for (size_t s : members[gi]) {
        const SolveState& S = *solves[s];
        for (const Tile& t : tiles_of(*S.T, g.lds)) all.push_back(Tile{int32_t(s), t.blk, t.start, t.end, t.aux});
      }
      g.tiles = xcd_split(all, tiles_per_wg(g.lds, g.D, g.K), tiles_per_wave(g.lds, g.D, g.K));
      if (g.lds.kernel == KERNEL_FIBER) {
        int32_t slots = 0, ents = 0;
        for (size_t s : members[gi]) {
          slots = std::max(slots, solves[s]->T->tplan[fp32].max_slots);
          ents = std::max(ents, solves[s]->T->tplan[fp32].max_ents);
        }
        g.lds = lds_layout(g.D, g.K, fp32, slots, ents);
      }
      upload(g.dtiles, g.tiles);
    }
    upload(ddescs, descs);
    assign_lanes();
    build_fused();
  }

  void build_fused() {
    schedule = fp32 ? resolve_schedule(schedule_own) : GSG_SCHEDULE_STREAMS;
    fused[0] = FusedArgs{};
    fused[1] = FusedArgs{};
    unfused.clear();
    fused_pass_a = FUSED_SUM;
    if (schedule != GSG_SCHEDULE_FUSED) return;
    std::vector<size_t> order(groups.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    auto cost = [&](size_t i) { return double(groups[i].dofs) * groups[i].D * groups[i].K; };
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost(x) > cost(y); });
    auto add = [&](FusedArgs& f, const Group& g, int op) {
      FusedSeg& q = f.seg[f.nseg++];
      q.tiles = g.dtiles.as<Tile>();
      q.ntiles = int32_t(g.tiles.size());
      q.wg_end = f.items + int32_t(g.tiles.size() / tiles_per_wg(g.lds, g.D, g.K));
      q.body = int16_t(fused_body(g.lds.kernel, g.D, g.K));
      q.op = int16_t(op);
      f.items = q.wg_end;
    };
    for (size_t gi : order) {
      const Group& g = groups[gi];
      const bool wave = g.sys == SYS_WAVE;
      if (!fused_ok(g.lds.kernel, g.D, g.K) || fused[0].nseg >= FUSED_MAXSEG || (wave && fused[1].nseg >= FUSED_MAXSEG)) {
        unfused.push_back(gi);
        continue;
      }
      add(fused[0], g, wave ? OP_RK_WAVE_GRAD : OP_RK_SUM);
      if (wave) {
        add(fused[1], g, OP_RK_WAVE_SUM);
        fused_pass_a = FUSED_A;
      }
    }
    fused[0].solves = fused[1].solves = ddescs.as<SolveDesc>();
    for (int p = 0; p < 2; ++p) {
      if (!fused[p].nseg) continue;
      // each class's items in runs of 8 (run j of a class: its workgroups 8 j .. 8 j + 7,
      // one per XCD), the runs of all classes spread evenly over the launch: run j of a
      // class with R runs sits at (j + 1/2) / R
      std::vector<std::pair<double, std::pair<int, int>>> runs;
      for (int q = 0; q < fused[p].nseg; ++q) {
        const int wgs = fused[p].seg[q].wg_end - (q ? fused[p].seg[q - 1].wg_end : 0);
        if (wgs % NXCD) throw std::logic_error("fused class items must come in whole XCD runs");
        const int R = wgs / NXCD;
        for (int j = 0; j < R; ++j) runs.push_back({(j + 0.5) / R, {q, j}});
      }
      std::stable_sort(runs.begin(), runs.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      std::vector<int32_t> ord;
      ord.reserve(size_t(fused[p].items));
      for (const auto& r : runs)
        for (int x = 0; x < NXCD; ++x) ord.push_back((r.second.first << 24) | (r.second.second * NXCD + x));
      if (int64_t(ord.size()) != fused[p].items || fused[p].items >= (1 << 24))
        throw std::logic_error("fused item order does not cover the items");
      upload(fused_order[p], ord);
      fused[p].order = fused_order[p].as<int32_t>();
    }
  }

  // change the schedule of a built batch (the captured steps are recorded anew)
  void set_schedule(int s) {
    schedule_own = s;
    reset_graph();
    build_fused();
  }

  // stop advancing solve s of the last build (its launches then skip it); call with the
  // stream idle.  The captured step graphs stay valid: they read the descriptors.
  void deactivate(size_t s) {
    if (s >= descs.size() || !descs[s].active) return;
    descs[s].active = 0;
    CaptureGuard cg(g_capture_mu);
    HIPC(hipMemcpy(reinterpret_cast<unsigned char*>(ddescs.p) + s * sizeof(SolveDesc) + offsetof(SolveDesc, active),
                   &descs[s].active, sizeof(int32_t), hipMemcpyHostToDevice));
  }

  hipEvent_t next_event() {
    if (evused == evpool.size()) {
      hipEvent_t e;
      {
        CaptureGuard cg(g_capture_mu);
        HIPC(hipEventCreate(&e));
      }
      evpool.push_back(e);
    }
    return evpool[evused++];
  }

  void launch(const Group& g, int op, int stage, hipStream_t stream) {
    LaunchArgs a{};
    a.solves = ddescs.as<SolveDesc>();
    a.tiles = g.dtiles.as<Tile>();
    a.ntiles = int32_t(g.tiles.size());
    a.op = op;
    a.stage = stage;
    set_lds(a, g.lds);
    if (!profile) {
      HIPC(launch_op(g.D, g.K, fp32, a, stream));
      return;
    }
    const size_t e0 = evused;
    HIPC(hipEventRecord(next_event(), stream));
    HIPC(launch_op(g.D, g.K, fp32, a, stream));
    HIPC(hipEventRecord(next_event(), stream));
    pend.push_back(Pending{g.D, g.K, g.sys, g.tiled(), g.dofs, double(g.dofs) * stage_bytes(g.sys, op, stage, fp32, g.D), e0});
  }

  // GSG_STREAMS=1 runs every launch on the batch stream (for kernel-trace profiles whose
  // per-kernel durations must not include other classes' concurrent kernels)
  static bool lanes_enabled() {
    static const bool on = [] {
      const char* e = std::getenv("GSG_STREAMS");
      return !(e && std::string(e) == "1");
    }();
    return on;
  }

  // RK4 steps of every group captured per build() (the fork to the lanes, the stage
  // launches, the joins) and replayed
  StepGraphs graphs;
  void reset_graph() { graphs.reset(); }

  void steps(int64_t n) {
    if (n <= 0) return;
    if (profile) {
      enqueue(n);
      return;
    }
    if (groups.size() > 1 && lanes_enabled() && schedule != GSG_SCHEDULE_FUSED)
      init_lanes();  // no stream creation inside a capture
    graphs.run(n, stream, true, [&](int64_t k, hipStream_t) { enqueue(k); });
  }

  void enqueue(int64_t n) {
    if (schedule == GSG_SCHEDULE_FUSED && !profile) {
      for (int64_t st = 0; st < n; ++st)
        for (int stage = 1; stage <= 4; ++stage) {
          if (fused[0].nseg) HIPC(launch_fused(fused_pass_a, stage, fused[0], fused[0].items, stream));
          for (size_t gi : unfused)
            launch(groups[gi], groups[gi].sys == SYS_WAVE ? OP_RK_WAVE_GRAD : OP_RK_SUM, stage, stream);
          if (fused[1].nseg) HIPC(launch_fused(FUSED_B, stage, fused[1], fused[1].items, stream));
          for (size_t gi : unfused)
            if (groups[gi].sys == SYS_WAVE) launch(groups[gi], OP_RK_WAVE_SUM, stage, stream);
        }
      return;
    }
    const bool par = !profile && groups.size() > 1 && lanes_enabled();
    if (par) {
      init_lanes();
      HIPC(hipEventRecord(fork, stream));
      for (int l = 0; l < nlanes; ++l) HIPC(hipStreamWaitEvent(lanes[l], fork, 0));
    }
    for (int64_t st = 0; st < n; ++st)
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        const Group& g = groups[gi];
        hipStream_t s = par ? lanes[lane_of[gi]] : stream;
        for (int stage = 1; stage <= 4; ++stage) {
          if (g.sys == SYS_WAVE) {
            launch(g, OP_RK_WAVE_GRAD, stage, s);
            launch(g, OP_RK_WAVE_SUM, stage, s);
          } else {
            launch(g, OP_RK_SUM, stage, s);
          }
        }
      }
    if (par) {
      for (int l = 0; l < nlanes; ++l) {
        HIPC(hipEventRecord(joins[l], lanes[l]));
        HIPC(hipStreamWaitEvent(stream, joins[l], 0));
      }
    }
  }

  // after the stream is synchronised: fold the recorded launch times into stats
  void collect() {
    for (const Pending& p : pend) {
      float ms = 0.f;
      HIPC(hipEventElapsedTime(&ms, evpool[p.ev], evpool[p.ev + 1]));
      gsg_kernel_stat& s = stats[std::make_tuple(p.D, p.K, p.sys, p.tiled)];
      s.D = p.D;
      s.k = p.K;
      s.precision = fp32 ? GSG_PREC_FP32 : GSG_PREC_FP64;
      s.sys = p.sys;
      s.variant = p.tiled ? GSG_KERNEL_TILE : GSG_KERNEL_PLAN;
      s.launches += 1;
      s.total_ms += ms;
      s.algo_bytes += p.bytes;
      s.dofs = p.dofs;
    }
    pend.clear();
    evused = 0;
  }
};

}  // namespace gsg

struct gsg_batch {
  bool fp32 = false;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<std::unique_ptr<BatchSolve>> solves;
  bool built = false;
  Scheduler sched;
  std::vector<size_t> active;  // indices of solves in the current schedule
};

static void batch_reschedule(gsg_batch* b) {
  std::vector<SolveState*> act;
  b->active.clear();
  for (size_t i = 0; i < b->solves.size(); ++i)
    if (b->solves[i]->done < b->solves[i]->nsteps) {
      act.push_back(&b->solves[i]->st);
      b->active.push_back(i);
    }
  b->sched.build(act);
}

static void batch_build(gsg_batch* b) {
  if (b->built) return;
  for (auto& s : b->solves) {
    auto T = get_tables(s->d.D, s->d.k, s->d.n, s->d.scheme);
    s->st.init(T, s->d, b->fp32);
    s->st.desc.dt = s->dt;
    require(int64_t(s->y0.size()) == s->st.len, "initial state length does not match the grid");
    HIPC(hipMemcpyAsync(s->st.y.p, s->y0.data(), s->y0.size() * 8, hipMemcpyHostToDevice, b->stream));
    s->st.to_stage0(b->stream);
  }
  HIPC(hipStreamSynchronize(b->stream));
  b->sched.fp32 = b->fp32;
  b->sched.stream = b->stream;
  batch_reschedule(b);
  b->built = true;
}

// Advance all unfinished solves by up to `steps` steps, in phases so that no solve
// overshoots its own step count.
static void batch_advance(gsg_batch* b, int64_t steps) {
  batch_build(b);
  while (steps > 0 && !b->active.empty()) {
    int64_t phase = steps;
    for (size_t i : b->active) phase = std::min(phase, b->solves[i]->nsteps - b->solves[i]->done);
    b->sched.steps(phase);
    if (b->sched.profile) {
      HIPC(hipStreamSynchronize(b->stream));
      b->sched.collect();
    }
    steps -= phase;
    bool finished = false;
    for (size_t i : b->active) {
      b->solves[i]->done += phase;
      finished |= b->solves[i]->done >= b->solves[i]->nsteps;
    }
    if (finished) {
      HIPC(hipStreamSynchronize(b->stream));
      batch_reschedule(b);
    }
  }
}

extern "C" {

int gsg_batch_create(int32_t precision, gsg_batch** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    require(precision == GSG_PREC_FP64 || precision == GSG_PREC_FP32, "unknown precision");
    ensure_device();
    auto b = std::make_unique<gsg_batch>();
    b->fp32 = precision == GSG_PREC_FP32;
    HIPC(hipGetDevice(&b->device));
    {
      CaptureGuard cg(g_capture_mu);
      HIPC(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    }
    *out = b.release();
  });
}

int gsg_batch_add(gsg_batch* b, const gsg_op_desc* desc, const double* y0, int64_t len, double dt, int64_t nsteps,
                  int32_t* slot) {
  return guarded([&] {
    require(b && y0 && slot, "NULL argument");
    require(!b->built, "batch already started");
    check_desc(desc);
    require(desc->kind == GSG_OP_TRANSPORT || desc->kind == GSG_OP_WAVE, "batch solves are transport or wave");
    require(nsteps >= 0 && std::isfinite(dt), "bad dt/nsteps");
    const int64_t N = layout_size(desc->D, desc->k, desc->n, desc->scheme == GSG_SCHEME_FULL);
    require(len == (desc->kind == GSG_OP_WAVE ? 2 * N : N), "len does not match the grid");
    auto s = std::make_unique<BatchSolve>();
    s->d = *desc;
    s->y0.assign(y0, y0 + len);
    s->dt = dt;
    s->nsteps = nsteps;
    *slot = int32_t(b->solves.size());
    b->solves.push_back(std::move(s));
  });
}

int gsg_batch_run(gsg_batch* b) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    batch_advance(b, INT64_MAX);
    HIPC(hipStreamSynchronize(b->stream));
  });
}

int gsg_batch_step(gsg_batch* b, int64_t steps) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(steps >= 0, "steps must be >= 0");
    batch_advance(b, steps);
    HIPC(hipStreamSynchronize(b->stream));
  });
}

int gsg_batch_get(gsg_batch* b, int32_t slot, double* y, int64_t len) {
  return guarded([&] {
    require(b && y, "NULL argument");
    require(slot >= 0 && size_t(slot) < b->solves.size(), "slot out of range");
    BatchSolve& s = *b->solves[slot];
    require(len == int64_t(s.y0.size()), "len does not match the solve");
    if (!b->built) {
      std::copy(s.y0.begin(), s.y0.end(), y);
      return;
    }
    HIPC(hipMemcpyAsync(y, s.st.y.p, size_t(len) * 8, hipMemcpyDeviceToHost, b->stream));
    HIPC(hipStreamSynchronize(b->stream));
  });
}

int gsg_batch_schedule(gsg_batch* b, int32_t schedule) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(schedule == GSG_SCHEDULE_AUTO || schedule == GSG_SCHEDULE_STREAMS || schedule == GSG_SCHEDULE_FUSED,
            "unknown schedule");
    if (b->built) {
      HIPC(hipStreamSynchronize(b->stream));
      b->sched.set_schedule(schedule);
    } else {
      b->sched.schedule_own = schedule;
    }
  });
}

int gsg_tune_lanes(int32_t lanes) {
  return guarded([&] {
    require(lanes == 0 || (lanes >= 1 && lanes <= 8), "lanes must be in 1..8 (0: the tuned count)");
    std::call_once(g_tune_loaded, load_tune_file);
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_lanes = lanes == 0 ? -1 : lanes;
  });
}

int gsg_tune_schedule(int32_t schedule) {
  return guarded([&] {
    require(schedule == GSG_SCHEDULE_AUTO || schedule == GSG_SCHEDULE_STREAMS || schedule == GSG_SCHEDULE_FUSED,
            "unknown schedule");
    std::call_once(g_tune_loaded, load_tune_file);
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_schedule = schedule == GSG_SCHEDULE_AUTO ? -1 : schedule;
  });
}

int gsg_batch_profile(gsg_batch* b, int32_t enable) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    b->sched.profile = enable != 0;
  });
}

int gsg_batch_kernel_stats(gsg_batch* b, gsg_kernel_stat* out, int32_t max, int32_t* count) {
  return guarded([&] {
    require(b && count, "NULL argument");
    int32_t c = 0;
    for (auto& kv : b->sched.stats) {
      if (out && c < max) out[c] = kv.second;
      ++c;
    }
    *count = c;
  });
}

int64_t gsg_batch_dofs(const gsg_batch* b) {
  if (!b) return GSG_EINVAL;
  int64_t t = 0;
  for (auto& s : b->solves) t += int64_t(s->y0.size());
  return t;
}

int gsg_batch_destroy(gsg_batch* b) {
  if (!b) return GSG_OK;
  return guarded([&] {
    if (b->stream) {
      (void)hipStreamSynchronize(b->stream);
      {
        CaptureGuard cg(g_capture_mu);
        (void)hipStreamDestroy(b->stream);
      }
    }
    delete b;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// continuous-batching server

namespace gsg {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  gsg_op_desc d{};
  std::vector<double> y0, result;
  double dt = 0;
  int64_t nsteps = 0, done = 0, len = 0;
  size_t bytes = 0;
  // the grid's operator tables (shared by every solve of the grid, built at submission)
  // and their device bytes, charged to the budget by the first admitted solve of the grid
  std::shared_ptr<Tables> tab;
  size_t tbytes = 0;
  int32_t status = GSG_REQ_QUEUED;
  std::string error;
  int32_t ecode = GSG_ESTATE;  // the code gsg_server_wait returns for a failed request
  Clock::time_point t_submit, t_done;
  std::unique_ptr<SolveState> st;
};

}  // namespace gsg

struct gsg_server {
  gsg_server_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;
  std::vector<Request*> resident;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  int64_t next_id = 1;
  int64_t budget = 0, used = 0;
  std::set<const Tables*> charged;  // grids whose tables count in `used`
  // admission is first fit in arrival order; the oldest waiting request is passed by
  // later ones for at most HOL_TICKS admission passes, then admission waits for it
  static constexpr int HOL_TICKS = 8;
  int64_t head_id = 0;
  int head_passes = 0;
  bool stop = false;
  gsg_server_stats stats{};
  Scheduler sched;
  // the solves of the scheduler's last build, in its order; finished ones stay there,
  // inactive, and their device state waits in `retired` until the next rebuild, so a
  // completion costs no rebuild of the batch
  std::vector<Request*> built;
  std::vector<std::pair<std::unique_ptr<SolveState>, int64_t>> retired;  // (state, bytes)
  bool dirty = true;
  // DoFs of the last build and of its solves that have finished since: once the finished
  // ones exceed a tenth, the next tick rebuilds (their skipped launches cost more than
  // a rebuild)
  int64_t built_dofs = 0, idle_dofs = 0;
  std::thread worker;
  std::string fatal;
  // entry-point calls in progress (submit, poll, wait, stats): stop frees the server only
  // once they have all returned
  std::atomic<int64_t> calls{0};
};

namespace gsg {
struct ServerCall {
  gsg_server* S;
  explicit ServerCall(gsg_server* s) : S(s) {
    if (S) S->calls.fetch_add(1);
  }
  ~ServerCall() {
    if (S) S->calls.fetch_sub(1);
  }
};
}  // namespace gsg

static void server_loop(gsg_server* S) {
  (void)hipSetDevice(S->device);
  for (;;) {
    std::vector<Request*> admit;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->stop || !S->pending.empty() || !S->resident.empty(); });
      if (S->stop) break;
      // retired solves free their device memory before requests waiting for it are
      // admitted (the stream is idle: every tick ends with a synchronisation)
      if (!S->retired.empty() && !S->pending.empty()) {
        for (auto& rs : S->retired) S->used -= rs.second;
        S->retired.clear();
        S->dirty = true;
      }
      // admission by device footprint (the solve's state, and its grid's tables when no
      // admitted solve has charged them yet): first fit in arrival order, except that
      // once the oldest request has waited HOL_TICKS passes nothing passes it
      for (auto it = S->pending.begin(); it != S->pending.end();) {
        Request* r = *it;
        const bool cap_ok = S->cfg.max_resident <= 0 || int64_t(S->resident.size() + admit.size()) < S->cfg.max_resident;
        if (!cap_ok) break;
        const int64_t need = int64_t(r->bytes) + (S->charged.count(r->tab.get()) ? 0 : int64_t(r->tbytes));
        if (S->used + need <= S->budget) {
          S->used += need;
          S->charged.insert(r->tab.get());
          admit.push_back(r);
          it = S->pending.erase(it);
          continue;
        }
        if (it == S->pending.begin()) {
          if (S->head_id != r->id) {
            S->head_id = r->id;
            S->head_passes = 0;
          }
          if (++S->head_passes > gsg_server::HOL_TICKS) break;
        }
        ++it;
      }
      // nothing resident and the oldest request still does not fit: the tables of the
      // grids served before hold the difference, and waiting frees nothing
      if (admit.empty() && S->resident.empty() && !S->pending.empty()) {
        Request* r = S->pending.front();
        S->pending.pop_front();
        r->status = GSG_REQ_FAILED;
        r->error = "request and its tables exceed the HBM budget left by the operator tables of earlier grids";
        r->ecode = GSG_EFULL;
        r->t_done = Clock::now();
        S->cv_done.notify_all();
        continue;
      }
    }
    if (!admit.empty()) S->dirty = true;
    for (Request* r : admit) {
      try {
        auto T = r->tab;
        r->st = std::make_unique<SolveState>();
        r->st->init(T, r->d, S->cfg.precision == GSG_PREC_FP32);
        r->st->desc.dt = r->dt;
        HIPC(hipMemcpyAsync(r->st->y.p, r->y0.data(), r->y0.size() * 8, hipMemcpyHostToDevice, S->stream));
        r->st->to_stage0(S->stream);
        std::lock_guard<std::mutex> g(S->mu);
        r->status = GSG_REQ_RUNNING;
        S->resident.push_back(r);
        S->stats.admitted++;
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> g(S->mu);
        r->status = GSG_REQ_FAILED;
        r->error = e.what();
        r->st.reset();
        S->used -= int64_t(r->bytes);
        r->t_done = Clock::now();
        S->cv_done.notify_all();
      }
    }
    std::vector<Request*> res;
    {
      std::lock_guard<std::mutex> g(S->mu);
      res = S->resident;
    }
    if (res.empty()) continue;
    int64_t tick = std::max<int32_t>(1, S->cfg.steps_per_tick);
    for (Request* r : res) tick = std::min(tick, r->nsteps - r->done);
    const auto t0 = Clock::now();
    try {  // nothing may escape the worker thread: a failure fails the resident requests
      if (S->dirty) {
        std::vector<SolveState*> ss;
        for (Request* r : res) ss.push_back(r->st.get());
        S->sched.build(ss);
        S->built = res;
        S->dirty = false;
        S->built_dofs = 0;
        for (Request* r : res) S->built_dofs += r->len;
        S->idle_dofs = 0;
        if (!S->retired.empty()) {  // no launch refers to them now
          {
            std::lock_guard<std::mutex> g(S->mu);
            for (auto& rs : S->retired) S->used -= rs.second;
          }
          S->retired.clear();
        }
      }
      if (tick > 0) S->sched.steps(tick);
      HIPC(hipStreamSynchronize(S->stream));
    } catch (const std::exception& e) {
      std::lock_guard<std::mutex> g(S->mu);
      S->fatal = e.what();
      for (Request* r : res) {
        r->status = GSG_REQ_FAILED;
        r->error = e.what();
        r->t_done = Clock::now();
        S->retired.emplace_back(std::move(r->st), int64_t(r->bytes));
      }
      S->resident.clear();
      S->built.clear();
      S->dirty = true;
      S->cv_done.notify_all();
      continue;
    }
    const auto t1 = Clock::now();
    std::vector<Request*> fin;
    for (Request* r : res) {
      r->done += tick;
      if (r->done >= r->nsteps) fin.push_back(r);
    }
    for (Request* r : fin) {
      r->result.resize(r->len);
      hipError_t e;
      {
        CaptureGuard cg(g_capture_mu);
        e = hipMemcpy(r->result.data(), r->st->y.p, size_t(r->len) * 8, hipMemcpyDeviceToHost);
      }
      bool finite = true;
      for (double v : r->result) finite &= std::isfinite(v);
      // the batch keeps its launches: the solve turns inactive, its state is freed at
      // the next rebuild
      const auto bi = std::find(S->built.begin(), S->built.end(), r);
      try {
        if (bi != S->built.end()) S->sched.deactivate(size_t(bi - S->built.begin()));
      } catch (const std::exception&) {
        S->dirty = true;  // rebuild instead
      }
      if (bi != S->built.end()) *bi = nullptr;
      S->idle_dofs += r->len;
      if (10 * S->idle_dofs > S->built_dofs) S->dirty = true;
      std::lock_guard<std::mutex> g(S->mu);
      S->retired.emplace_back(std::move(r->st), int64_t(r->bytes));
      r->t_done = Clock::now();
      r->status = (e == hipSuccess && finite) ? GSG_REQ_DONE : GSG_REQ_FAILED;
      if (e != hipSuccess) r->error = hipGetErrorString(e);
      else if (!finite) {
        r->error = "non-finite state";
        r->ecode = GSG_ENONFINITE;
      }
      S->resident.erase(std::find(S->resident.begin(), S->resident.end(), r));
      S->stats.completed++;
      S->stats.dof_updates += double(r->len) * double(r->nsteps);
    }
    {
      std::lock_guard<std::mutex> g(S->mu);
      S->stats.ticks++;
      S->stats.busy_seconds += std::chrono::duration<double>(t1 - t0).count();
      if (!fin.empty()) S->cv_done.notify_all();
    }
  }
}

extern "C" {

int gsg_server_start(const gsg_server_config* cfg, gsg_server** out) {
  return guarded([&] {
    require(cfg && out, "NULL argument");
    *out = nullptr;
    require(cfg->precision == GSG_PREC_FP64 || cfg->precision == GSG_PREC_FP32, "unknown precision");
    ensure_device();
    auto S = std::make_unique<gsg_server>();
    S->cfg = *cfg;
    if (S->cfg.steps_per_tick <= 0) S->cfg.steps_per_tick = 16;
    HIPC(hipGetDevice(&S->device));
    size_t fr = 0, tot = 0;
    HIPC(hipMemGetInfo(&fr, &tot));
    S->budget = cfg->hbm_budget_bytes > 0 ? cfg->hbm_budget_bytes : int64_t(double(fr) * 0.75);
    {
      CaptureGuard cg(g_capture_mu);
      HIPC(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
    }
    S->sched.fp32 = cfg->precision == GSG_PREC_FP32;
    S->sched.stream = S->stream;
    gsg_server* raw = S.get();
    S->worker = std::thread(server_loop, raw);
    *out = S.release();
  });
}

int gsg_server_submit(gsg_server* S, const gsg_op_desc* desc, const double* y0, int64_t len, double dt,
                      int64_t nsteps, int64_t* id) {
  ServerCall call(S);
  return guarded([&] {
    require(S && y0 && id, "NULL argument");
    check_desc(desc);
    require(desc->kind == GSG_OP_TRANSPORT || desc->kind == GSG_OP_WAVE, "served solves are transport or wave");
    require(nsteps >= 1 && std::isfinite(dt), "bad dt/nsteps");
    const int64_t N = layout_size(desc->D, desc->k, desc->n, desc->scheme == GSG_SCHEME_FULL);
    require(len == (desc->kind == GSG_OP_WAVE ? 2 * N : N), "len does not match the grid");
    auto r = std::make_unique<Request>();
    r->d = *desc;
    r->y0.assign(y0, y0 + len);
    r->dt = dt;
    r->nsteps = nsteps;
    r->len = len;
    r->bytes = SolveState::footprint(desc->kind, desc->D, N, S->cfg.precision == GSG_PREC_FP32);
    {
      DeviceScope ds(S->device);
      r->tab = get_tables(desc->D, desc->k, desc->n, desc->scheme);
    }
    r->tbytes = r->tab->device_bytes();
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(S->mu);
    if (S->stop) throw Error(GSG_ESTATE, "server stopped");
    const int64_t need = int64_t(r->bytes) + (S->charged.count(r->tab.get()) ? 0 : int64_t(r->tbytes));
    if (need > S->budget) throw Error(GSG_EFULL, "request exceeds the server's HBM budget");
    r->id = S->next_id++;
    *id = r->id;
    S->pending.push_back(r.get());
    S->reqs[r->id] = std::move(r);
    S->stats.submitted++;
    S->cv_work.notify_one();
  });
}

int gsg_server_poll(gsg_server* S, int64_t id, int32_t* status) {
  ServerCall call(S);
  return guarded([&] {
    require(S && status, "NULL argument");
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) throw Error(GSG_ESTATE, "unknown request id");
    *status = it->second->status;
  });
}

int gsg_server_wait(gsg_server* S, int64_t id, double* y, int64_t len, int64_t* latency_ns) {
  ServerCall call(S);
  return guarded([&] {
    require(S != nullptr, "NULL server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) throw Error(GSG_ESTATE, "unknown request id");
    Request* r = it->second.get();
    // a request leaves the queues (and the worker's hands) before it turns DONE or
    // FAILED, so erasing it below frees nothing the worker still uses; stop fails the
    // requests it leaves behind only after the worker has ended
    S->cv_done.wait(lk, [&] {
      // another waiter on the same id may have released it (and the map may have
      // rehashed since the lookup above: look it up again)
      auto jt = S->reqs.find(id);
      return jt == S->reqs.end() || jt->second.get() != r || r->status == GSG_REQ_DONE ||
             r->status == GSG_REQ_FAILED;
    });
    it = S->reqs.find(id);
    if (it == S->reqs.end() || it->second.get() != r) throw Error(GSG_ESTATE, "request released by another wait");
    if (r->status != GSG_REQ_DONE) {
      const std::string msg = r->error.empty() ? "server stopped" : r->error;
      const int code = r->error.empty() ? GSG_ESTATE : r->ecode;
      S->reqs.erase(it);
      S->cv_done.notify_all();  // other waiters on this id
      throw Error(code, msg);
    }
    if (latency_ns)
      *latency_ns = std::chrono::duration_cast<std::chrono::nanoseconds>(r->t_done - r->t_submit).count();
    if (y) {
      require(len == r->len, "len does not match the request");
      std::copy(r->result.begin(), r->result.end(), y);
    }
    S->reqs.erase(it);
    S->cv_done.notify_all();  // other waiters on this id
  });
}

int gsg_server_stats_get(gsg_server* S, gsg_server_stats* st) {
  ServerCall call(S);
  return guarded([&] {
    require(S && st, "NULL argument");
    std::lock_guard<std::mutex> g(S->mu);
    *st = S->stats;
    st->queued = int64_t(S->pending.size());
    st->resident = int64_t(S->resident.size());
    st->resident_bytes = S->used;
  });
}

int gsg_server_stop(gsg_server* S) {
  if (!S) return GSG_OK;
  return guarded([&] {
    {
      std::lock_guard<std::mutex> g(S->mu);
      S->stop = true;  // later submissions fail
      S->cv_work.notify_all();
    }
    // the worker finishes its tick and ends; it touches no request afterwards
    if (S->worker.joinable()) S->worker.join();
    {
      std::lock_guard<std::mutex> g(S->mu);
      for (auto& kv : S->reqs) {
        Request* r = kv.second.get();
        if (r->status == GSG_REQ_QUEUED || r->status == GSG_REQ_RUNNING) {
          r->status = GSG_REQ_FAILED;
          r->error = "server stopped";
          r->t_done = Clock::now();
        }
      }
      S->pending.clear();
      S->resident.clear();
      S->built.clear();
      S->cv_done.notify_all();
    }
    // blocked waiters wake, see their request failed and return; no call may still be
    // inside the server when it is freed
    while (S->calls.load() != 0) {
      {
        std::lock_guard<std::mutex> g(S->mu);
        S->cv_done.notify_all();
      }
      std::this_thread::sleep_for(std::chrono::milliseconds(1));
    }
    for (auto& kv : S->reqs) kv.second->st.reset();
    S->retired.clear();
    if (S->stream) {
      (void)hipStreamSynchronize(S->stream);
      {
        CaptureGuard cg(g_capture_mu);
        (void)hipStreamDestroy(S->stream);
      }
    }
    delete S;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// projection of initial data on the device (vcoeffs_DG, dg_vmethods.jl:149-179)

struct gsg_proj {
  int D = 0, k = 0, n = 0, q = 0, P = 0, full = 0, device = 0;
  int64_t N = 0, ncells = 0, ppc = 0, kD = 0;
  DevMem cell_blk, blk_cell0, blk_lev, t, wv;
  DevMem scratch[2];  // pass outputs, even and odd passes
  ProjArgs args{};
};

namespace gsg {
// elements one pass writes per cell: R_t K with R_t = P^(D-1-t) K^t
static int64_t proj_pass_out(const gsg_proj* p, int t) {
  int64_t r = p->k;
  for (int i = 0; i < p->D - 1 - t; ++i) r *= p->P;
  for (int i = 0; i < t; ++i) r *= p->k;
  return r;
}
static void proj_range(const gsg_proj* p, int64_t cell0, int64_t ncell) {
  require(cell0 >= 0 && ncell >= 1 && cell0 + ncell <= p->ncells, "cell range outside the grid");
}
}  // namespace gsg

extern "C" {

int gsg_proj_create(int32_t D, int32_t k, int32_t n, int32_t scheme, int32_t q, gsg_proj** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    require(D >= 1 && D <= MAXD, "D must be in 1..8");
    require(k >= 1 && k <= KMAX, "k must be in 1..10");
    require(n >= 0 && n <= 20, "n must be in 0..20");
    require(q >= 1 && q <= 32, "q must be in 1..32");
    require(scheme == GSG_SCHEME_SPARSE || scheme == GSG_SCHEME_FULL, "unknown scheme");
    int64_t ppc = 1;
    for (int d = 0; d < D; ++d) {
      ppc *= 2 * q;
      require(ppc < (int64_t(1) << 31), "(2q)^D points per cell exceed 2^31: lower q");
    }
    ensure_device();
    auto p = std::make_unique<gsg_proj>();
    p->D = D; p->k = k; p->n = n; p->q = q; p->P = 2 * q; p->full = scheme == GSG_SCHEME_FULL;
    p->ppc = ppc;
    HIPC(hipGetDevice(&p->device));
    const Layout lt = build_layout(D, k, n, p->full);
    p->N = lt.N;
    p->kD = 1;
    for (int d = 0; d < D; ++d) p->kD *= k;
    p->ncells = lt.N / p->kD;
    require(p->ncells < (int64_t(1) << 31), "grid too large");
    std::vector<int32_t> cell_blk(size_t(p->ncells)), cell0(size_t(lt.nblk) + 1);
    for (int b = 0; b <= lt.nblk; ++b) cell0[b] = int32_t(lt.blk_off[b] / p->kD);
    for (int b = 0; b < lt.nblk; ++b) std::fill(cell_blk.begin() + cell0[b], cell_blk.begin() + cell0[b + 1], b);
    std::vector<double> nodes, wv;
    projection_tables(k, n, q, nodes, wv);
    upload(p->cell_blk, cell_blk);
    upload(p->blk_cell0, cell0);
    upload(p->blk_lev, lt.blk_lev);
    upload(p->t, nodes);
    upload(p->wv, wv);
    p->args = ProjArgs{p->cell_blk.as<int32_t>(), p->blk_cell0.as<int32_t>(), p->blk_lev.as<int32_t>(), D, k, q,
                       2 * q, ppc};
    *out = p.release();
  });
}

int64_t gsg_proj_len(const gsg_proj* p) { return p ? p->N : GSG_EINVAL; }
int64_t gsg_proj_cells(const gsg_proj* p) { return p ? p->ncells : GSG_EINVAL; }
int64_t gsg_proj_points_per_cell(const gsg_proj* p) { return p ? p->ppc : GSG_EINVAL; }

int gsg_proj_nodes(gsg_proj* p, int64_t cell0, int64_t ncell, double* X, int64_t ld, void* stream) {
  return guarded([&] {
    require(p && X, "NULL argument");
    proj_range(p, cell0, ncell);
    require(ld >= ncell * p->ppc, "ld must be >= ncell * points_per_cell");
    require((reinterpret_cast<uintptr_t>(X) & 7) == 0, "X must be 8-byte aligned");
    HIPC(launch_proj_nodes(p->args, p->t.as<double>(), cell0, ncell, X, ld, (hipStream_t)stream));
  });
}

int gsg_proj_nodes_1d(gsg_proj* p, int64_t cell0, int64_t ncell, double* X1, int64_t len, void* stream) {
  return guarded([&] {
    require(p && X1, "NULL argument");
    proj_range(p, cell0, ncell);
    require(len == int64_t(p->D) * ncell * p->P, "len must be D * ncell * 2q");
    require((reinterpret_cast<uintptr_t>(X1) & 7) == 0, "X1 must be 8-byte aligned");
    HIPC(launch_proj_nodes_1d(p->args, p->t.as<double>(), cell0, ncell, X1, (hipStream_t)stream));
  });
}

int gsg_proj_contract(gsg_proj* p, int64_t cell0, int64_t ncell, const double* fvals, int64_t fvals_len,
                      double* out, int64_t out_len, void* stream) {
  return guarded([&] {
    require(p && fvals && out, "NULL argument");
    proj_range(p, cell0, ncell);
    require(fvals_len == ncell * p->ppc, "fvals_len must be ncell * points_per_cell");
    require(out_len == p->N, "out_len must be get_size(D, k, n)");
    require((reinterpret_cast<uintptr_t>(fvals) & 7) == 0 && (reinterpret_cast<uintptr_t>(out) & 7) == 0,
            "fvals and out must be 8-byte aligned");
    const hipStream_t s = (hipStream_t)stream;
    for (int t = 0; t + 1 < p->D; ++t) {  // every pass but the last writes scratch[t & 1]
      const size_t need = size_t(ncell * proj_pass_out(p, t)) * 8;
      if (p->scratch[t & 1].bytes < need) {
        HIPC(hipDeviceSynchronize());  // earlier passes may still read the old buffer
        p->scratch[t & 1].alloc(need);
      }
    }
    const double* in = fvals;
    for (int t = 0; t < p->D; ++t) {
      const int d = p->D - 1 - t;
      const int64_t R = proj_pass_out(p, t) / p->k;
      double* dst = t + 1 == p->D ? out + cell0 * p->kD : p->scratch[t & 1].as<double>();
      HIPC(launch_proj_pass(p->args, p->wv.as<double>(), d, cell0, ncell, R, in, dst, s));
      in = dst;
    }
  });
}

int gsg_proj_tensor(gsg_proj* p, const double* const* vecs, double* out, int64_t out_len, void* stream) {
  return guarded([&] {
    require(p && vecs && out, "NULL argument");
    require(out_len == p->N, "out_len must be get_size(D, k, n)");
    TensorVecs tv{};
    for (int d = 0; d < p->D; ++d) {
      require(vecs[d] != nullptr, "NULL 1-D vector");
      tv.v[d] = vecs[d];
    }
    HIPC(launch_tensor(p->args, tv, p->N, out, (hipStream_t)stream));
  });
}

int gsg_proj_destroy(gsg_proj* p) {
  return guarded([&] {
    if (!p) return;
    (void)hipDeviceSynchronize();
    delete p;
  });
}

}  // extern "C"
