// Host side of the MI355X coupled-cluster amplitude engine: IR loader,
// planner (offset tables for every block contraction), resident device
// state, and the C ABI declared in include/ccq_hip.h.
//
// The contraction program is produced offline by ccq_amd/gen (spin-orbital
// equations -> spin-integrated block contractions) and embedded as text.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ccq_hip.h"
#include "kernels.h"
#include "restart_io.h"

#include "ir_ccsd_rhf.inc"
#include "ir_ccsd_uhf.inc"
#include "ir_ccsdt_rhf.inc"
#include "ir_ccsdt_uhf.inc"
#include "ir_ccsdt_act_uhf.inc"
#include "ir_cct3_uhf.inc"
#if __has_include("ir_ccsdtq_uhf.inc") && __has_include("ir_ccsdtq_rhf.inc")
#include "ir_ccsdtq_rhf.inc"
#include "ir_ccsdtq_uhf.inc"
#define CCQ_HAVE_CCSDTQ 1
#endif

namespace {

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      throw Fail(std::string(#x) + ": " + hipGetErrorString(e_));                  \
  } while (0)

constexpr int kPartialBlocks = 1024;
constexpr int64_t kWsCap = 16 << 20;   // split-K workspace bound (doubles)

// ---------------------------------------------------------------------------
// device buffers

// Pinned host array for the small device-to-host copies of the partial sums
// (energy, residuum, DIIS Gram matrix): a copy into pageable memory goes
// through a staging buffer and costs several times as long.
struct PinnedVec {
  double* p = nullptr;
  size_t cap = 0;
  void resize(size_t n) {
    if (n <= cap) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&p), sizeof(double) * n, hipHostMallocDefault));
    cap = n;
  }
  double* data() { return p; }
  double& operator[](size_t i) { return p[i]; }
  ~PinnedVec() {
    if (p) (void)hipHostFree(p);
  }
};

struct DevBuf {
  double* p = nullptr;
  int64_t n = 0;
  void alloc(int64_t count) {
    n = count;
    if (count > 0) HIPCHK(hipMalloc(&p, sizeof(double) * count));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// ---------------------------------------------------------------------------
// tensors and ops of the loaded program

enum class TKind { Inter, View, Amp, Resid };

struct Tensor {
  std::string id;
  TKind kind;
  int rank = 0;
  int64_t dims[8] = {0};
  int64_t strides[8] = {0};
  double* ptr = nullptr;
  float* p32 = nullptr;      // float shadow of the FP32-storage sweeps
  std::string base;          // View: base array; Amp/Resid: block name
  std::string spec[8];       // View: per-index "oa", "vb", ...
  // View: where the block lives inside the full-range base array; the view is
  // gathered into a dense C-order copy (ptr/strides) after every upload
  double* src = nullptr;
  int64_t src_strides[8] = {0};
  // position pairs (p < q) in which the block is antisymmetric
  std::vector<std::pair<int, int>> anti;
};

struct Op {
  int out = -1, a = -1, b = -1;
  std::string lc, la, lb;
  double coef = 1.0;
  bool assign = false;
  bool gemm = false;
  int64_t M = 1, N = 1, K = 1;
  int64_t t_rowA = 0, t_kA = 0, t_kB = 0, t_colB = 0, t_rowC = 0, t_colC = 0;
  ccqk::GemmArgs plan{};      // launch plan (kernel, tile, split-K, affine strides)
  bool flat = false;          // add op over identical dense layouts
  // an add op folded into the epilogue of the GEMM that next updates its
  // output (fused: this add is not launched; fuse_a: the GEMM's addend)
  bool fused = false;
  int fuse_a = -1;
  // CCSDt: a triples residual product computed only on the P-space tuples
  // its masked update reads (one rectangle of the split output)
  bool restricted = false;
  // one of the parts an op was split into (P-space rectangles, irrep
  // blocks): parts of one op write disjoint elements, so they may share a
  // dependency level; gplan is the tile plan all parts share in a grouped launch
  int part_of = -1;
  bool has_gplan = false;
  ccqk::GemmArgs gplan{};
  // plan measured in FP32 storage (precision mode 2) for a product of the
  // sweep launched on its own; grouped and chained launches keep their plans
  bool has_plan32 = false;
  ccqk::GemmArgs plan32{};
  double dcoef = 0.0;
  int64_t t_rowD = 0, t_colD = 0;
  // antisymmetric pairs packed to their canonical half: a summed pair (k),
  // an output pair from A (m, mirrored rows rowC2) or from B (n, colC2)
  bool pack_k = false;
  int64_t t_rowC2 = -1, t_colC2 = -1;
  // chain of GEMMs into one output run as one launch: index into the
  // section's chains (the head is launched, the other members are skipped)
  int chain = -1;
  bool chain_head = false;
  double flops = 0, bytes = 0;
  std::string cls;           // profile class
};

struct UpdateStep {
  std::string block;
  int resid = -1;
  int nimg = 0;
  std::vector<int> perms;    // nimg x 8
  std::vector<double> signs;
  int dev_perm_off = 0;      // offset into the int table
  int dev_sign_off = 0;
  bool masked = false;       // active-space triples: update only the masked tuples
  int vneed = 0, vforbid = 0;
  bool left = false;         // left-CC step: H-bar denominators, |residual|^2 partials
  int sq_off = 0, sq_n = 0;  // this step's slice of the partial-sum buffer
  // canonical occupied / virtual halves of the block's tuples (int table offsets)
  int t_coc = -1, t_cvc = -1;
  int64_t n_oc = 0, n_vc = 0;
  int m3 = -1;               // triples correction: the M3 accumulator (resid = L3)
};

struct Section {
  std::string name;
  std::string prog = "sweep";   // program the section belongs to
  enum { Compute, Update, Copy, Cct3 } kind = Compute;
  std::vector<Op> ops;
  // launch order of a compute section: indices into ops, one launch each;
  // a launch of several indices is a grouped GEMM launch.  Empty: ops in
  // order, one launch each.
  std::vector<std::vector<int>> launches;
  std::vector<std::string> launch_cls;
  std::vector<int> launch_level;   // dependency level of each launch (non-decreasing)
  // chains: member op indices (program order), their orientation against the
  // first member (1: transposed) and the launch plan over the concatenated K
  std::vector<std::vector<int>> chains;
  std::vector<std::vector<int>> chain_flip;
  std::vector<ccqk::GemmArgs> chain_plan;
  UpdateStep upd;
  std::string dst, src;
};

struct ProfEntry {
  int64_t calls = 0;
  double ms = 0, flops = 0, bytes = 0;
};

}  // namespace

struct ccq_hip_ctx {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  // sizes
  int64_t froz = 0, occ_a = 0, occ_b = 0, orbs = 0;
  int64_t oa = 0, ob = 0, va = 0, vb = 0, nc = 0;
  int level = 2;
  bool rhf = false;
  double shift = 0.0;
  int diis = 5;
  // amplitude layout (ccq t_vec)
  std::map<std::string, int64_t> blk_off, blk_size;
  std::map<std::string, std::vector<int64_t>> blk_dims;
  int64_t t_size = 0;        // full-storage T1..T3 (ccq's pos(10)-1)
  int64_t dev_t = 0;         // doubles of t_vec kept on the device
  int64_t vec_size = 0;      // DIIS vector length
  int64_t ext_size = 0;      // ccq's t_vec (t_size, or pos(16)-1 for CCSDt)
  // active-space CCSDt: T3 is held in full storage with its Q-space tuples
  // zero; ccq's t_vec packs the P space into ten blocks
  bool active = false;
  int64_t act_occ = -1, act_unocc = -1;
  int64_t occ_start = 0;     // first active occupied index (both spins)
  int64_t vir_end[2] = {0, 0};   // active virtuals per spin (the lowest ones)
  struct PackPiece {         // one packed block, or one antisymmetric image of it
    int64_t rows, cols, packed_off;
    int64_t t_full_row, t_full_col, t_pk_row, t_pk_col;   // offsets into pk_tab
    double sign;
    bool image;              // unpack-only image (pack uses the identity pieces)
    int blk;                 // index into pack_block
  };
  std::vector<PackPiece> pack_pieces;
  int64_t* pk_tab = nullptr;
  DevBuf staging;            // ccq-layout t_vec on the device
  // device state
  DevBuf Vaa, Vab, Vbb, Fa, Fb, fdiag_a, fdiag_b, tvec, ring, partial;
  double* vbb_ptr = nullptr;   // Vbb, or Vaa when ccq's v_bb equals v_aa
  std::map<std::string, std::unique_ptr<DevBuf>> resid_buf;
  std::vector<std::unique_ptr<DevBuf>> inter_buf;
  int* tables32 = nullptr;   // offset tables of every op
  int64_t* gather_tab = nullptr;
  std::vector<std::unique_ptr<DevBuf>> view_buf;
  struct Gather { int t; int64_t rows, cols, t_src_row, t_src_col, t_dst_row, t_dst_col; };
  std::vector<Gather> gathers;
  DevBuf ws;                 // split-K workspace (main stream)
  // launches of one dependency level run concurrently on up to nstreams
  // streams (the main one and side[0..nstreams-2], default priority), each
  // with its own split-K workspace; CCQ_HIP_STREAMS=1 runs everything on the
  // main stream
  static constexpr int kMaxStreams = 4;
  int nstreams = kMaxStreams;
  hipStream_t side[kMaxStreams - 1] = {};
  DevBuf ws_side[kMaxStreams - 1];
  hipEvent_t fork_ev = nullptr, join_ev[kMaxStreams - 1] = {};
  double* cur_ws = nullptr;  // workspace of the stream being launched on
  int* itab = nullptr;
  double* dtab = nullptr;
  // program
  std::vector<Tensor> tensors;
  std::map<std::string, int> tindex;
  std::vector<Section> sections;
  bool ints_ready = false;
  // arithmetic of the LDS-tiled contractions of the sweep: 0 FP64 MFMA,
  // 1 FP32 MFMA (mixed precision; everything else stays FP64), 2 FP32 MFMA
  // with FP32 storage: every tensor of the sweep has a float shadow (made at
  // the first such sweep, zeroed), the sweep casts t_vec to its shadow
  // first, and its products, adds and updates read and write the shadows;
  // the updates still write the FP64 amplitudes.  Other programs run mode 2
  // as mode 1.
  int prec = 0;
  int run_prec = 0;          // mode of the program being run
  // the float copy of t_vec equals t_vec: set by a mode-2 sweep (which casts
  // every block it updates or copies), cleared on entry to every C-ABI call
  // that may write t_vec (all but the read-only ones, see guarded); a mode-2
  // sweep casts t_vec first only when it was not valid at its call
  bool t32_valid = false, t32_valid_at_call = false;
  bool f32_ready = false;
  bool views32_stale = true;   // views gathered again since their last cast
  float* tvec32 = nullptr;
  std::vector<float*> f32_bufs;
  // CC(t;3) tail (level 6): H-bar, left-CCSD and the triples correction run
  // as programs of a second IR on the same resident tensors
  bool tail = false;
  DevBuf lvec;               // ccq's l_vec: T1/T2 layout
  int64_t l_size = 0;
  double left_ecor = 0.0;    // E_cor in the left denominators
  bool left_mode = false;    // DIIS entry points act on l_vec
  DevBuf lsq, cpart;         // left |LH|^2 partials; correction partials
  int64_t lsq_n = 0;
  int64_t* d3tab = nullptr;  // cal_d3 gather tables
  struct Had { int out; const double* X; const double* Y; int64_t M, K, xr, xk, yr, yk; double scale; };
  std::vector<Had> hads;
  // profiling
  bool profile = false;
  std::map<std::string, ProfEntry> prof;
  PinnedVec host_partial;

  ~ccq_hip_ctx() {
    for (auto& kv : graphs) drop_graph(kv.second);
    for (float* p : f32_bufs) (void)hipFree(p);
    if (guess_params) (void)hipFree(guess_params);
    if (d3tab) (void)hipFree(d3tab);
    if (tables32) (void)hipFree(tables32);
    if (pk_tab) (void)hipFree(pk_tab);
    if (gather_tab) (void)hipFree(gather_tab);
    if (itab) (void)hipFree(itab);
    if (dtab) (void)hipFree(dtab);
    for (auto& e : join_ev)
      if (e) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    for (auto& st : side)
      if (st) (void)hipStreamDestroy(st);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // l1a, l3aab, m3abb ... have the shape of t1a, t3aab, t3abb
  static std::string dimkey(const std::string& base) {
    if (!base.empty() && (base[0] == 'l' || base[0] == 'm')) return "t" + base.substr(1);
    return base;
  }

  int64_t dimsym(const std::string& s) const {
    if (s == "oa") return oa;
    if (s == "ob") return ob;
    if (s == "va") return va;
    if (s == "vb") return vb;
    throw Fail("unknown dimension symbol " + s);
  }

  void layout() {
    // ccq's t_vec (src/utils/cc_utils.f90:479-559): occupied-first C order
    struct B {
      const char* name;
      std::vector<int64_t> d;
    };
    std::vector<B> blocks = {
        {"t1a", {oa, va}},
        {"t1b", {ob, vb}},
        {"t2aa", {oa, oa, va, va}},
        {"t2ab", {oa, ob, va, vb}},
        {"t2bb", {ob, ob, vb, vb}},
        {"t3aaa", {oa, oa, oa, va, va, va}},
        {"t3aab", {oa, oa, ob, va, va, vb}},
        {"t3abb", {oa, ob, ob, va, vb, vb}},
        {"t3bbb", {ob, ob, ob, vb, vb, vb}},
        // T4 is not part of ccq's t_vec (ccq keeps it in files,
        // src/utils/cc_utils.f90:89-145); here it is resident after t_vec
        {"t4aaaa", {oa, oa, oa, oa, va, va, va, va}},
        {"t4aaab", {oa, oa, oa, ob, va, va, va, vb}},
        {"t4aabb", {oa, oa, ob, ob, va, va, vb, vb}},
        {"t4abbb", {oa, ob, ob, ob, va, vb, vb, vb}},
        {"t4bbbb", {ob, ob, ob, ob, vb, vb, vb, vb}},
    };
    int64_t off = 0;
    for (auto& b : blocks) {
      int64_t sz = 1;
      for (auto x : b.d) sz *= x;
      blk_off[b.name] = off;
      blk_size[b.name] = sz;
      blk_dims[b.name] = b.d;
      off += sz;
      if (std::string(b.name) == "t3bbb") t_size = off;
    }
    vec_size = (level == 2) ? blk_off["t3aaa"] : t_size;
    dev_t = (level == 2) ? blk_off["t3aaa"] : (level == 4 ? off : t_size);
    ext_size = t_size;
    if (active) {
      plan_packing();
      vec_size = ext_size;
    }
  }

  // ccq's CCSDt t_vec: T1, T2, then ten P-space T3 blocks with the index
  // ranges of src/cc_opt_ccsdt/t2B0000_update_opt.f:48-57 (sizes
  // get_t_sizes_act, src/utils/cc_utils.f90:560-631).  Each range is given
  // per Fortran slot (fastest first) as (space, spin, part).
  void plan_packing() {
    enum Part { All, Act, Inact };
    struct Slot { char space; int spin; Part part; };
    struct PB { const char* block; Slot s[6]; };
    const PB pbs[10] = {
        {"t3aaa", {{'v', 0, All}, {'v', 0, All}, {'v', 0, Act}, {'o', 0, All}, {'o', 0, All}, {'o', 0, Act}}},
        {"t3aab", {{'v', 1, All}, {'v', 0, All}, {'v', 0, Act}, {'o', 1, All}, {'o', 0, All}, {'o', 0, Act}}},
        {"t3aab", {{'v', 1, Act}, {'v', 0, Inact}, {'v', 0, Inact}, {'o', 1, Act}, {'o', 0, Inact}, {'o', 0, Inact}}},
        {"t3aab", {{'v', 1, Act}, {'v', 0, Inact}, {'v', 0, Inact}, {'o', 1, All}, {'o', 0, All}, {'o', 0, Act}}},
        {"t3aab", {{'v', 1, All}, {'v', 0, All}, {'v', 0, Act}, {'o', 1, Act}, {'o', 0, Inact}, {'o', 0, Inact}}},
        {"t3abb", {{'v', 1, All}, {'v', 1, Act}, {'v', 0, All}, {'o', 1, All}, {'o', 1, Act}, {'o', 0, All}}},
        {"t3abb", {{'v', 1, Inact}, {'v', 1, Inact}, {'v', 0, Act}, {'o', 1, Inact}, {'o', 1, Inact}, {'o', 0, Act}}},
        {"t3abb", {{'v', 1, Inact}, {'v', 1, Inact}, {'v', 0, Act}, {'o', 1, All}, {'o', 1, Act}, {'o', 0, All}}},
        {"t3abb", {{'v', 1, All}, {'v', 1, Act}, {'v', 0, All}, {'o', 1, Inact}, {'o', 1, Inact}, {'o', 0, Act}}},
        {"t3bbb", {{'v', 1, All}, {'v', 1, All}, {'v', 1, Act}, {'o', 1, All}, {'o', 1, All}, {'o', 1, Act}}},
    };
    auto range = [&](const Slot& sl, int64_t& lo, int64_t& hi) {
      if (sl.space == 'o') {
        const int64_t n = sl.spin ? ob : oa;
        const int64_t st = std::min(occ_start, n);
        lo = sl.part == Act ? st : 0;
        hi = sl.part == Inact ? st : n;
      } else {
        const int64_t n = sl.spin ? vb : va;
        const int64_t e = std::min(vir_end[sl.spin], n);
        lo = sl.part == Inact ? e : 0;
        hi = sl.part == Act ? e : n;
      }
    };
    std::vector<int64_t> host;
    int64_t off = blk_off.at("t3aaa");
    pack_block.clear();
    for (const PB& pb : pbs) {
      pack_block.push_back(pb.block);
      const std::string name = pb.block;
      const auto& d = blk_dims.at(name);
      // C-order strides of the full block; Fortran slot s is axis 5 - s
      int64_t str[6];
      str[5] = 1;
      for (int a = 4; a >= 0; --a) str[a] = str[a + 1] * d[a + 1];
      int64_t lo[6], hi[6], len[6];
      int64_t count = 1;
      for (int sl = 0; sl < 6; ++sl) {
        range(pb.s[sl], lo[sl], hi[sl]);
        len[sl] = std::max<int64_t>(0, hi[sl] - lo[sl]);
        count *= len[sl];
      }
      // antisymmetric groups of the block (axes): same-spin occupied / virtual runs
      const int nb = name == "t3aaa" ? 0 : name == "t3aab" ? 1 : name == "t3abb" ? 2 : 3;
      const int na = 3 - nb;
      std::vector<std::vector<int>> groups;
      for (auto g : {std::vector<int>{0, 1, 2}, std::vector<int>{3, 4, 5}}) {
        std::vector<int> ga(g.begin(), g.begin() + na), gb(g.begin() + na, g.end());
        if (ga.size() > 1) groups.push_back(ga);
        if (gb.size() > 1) groups.push_back(gb);
      }
      // every combination of within-group permutations: image q[a] = p[perm[a]]
      std::vector<std::pair<std::vector<int>, double>> images;
      std::function<void(size_t, std::vector<int>, double)> gen = [&](size_t gi, std::vector<int> perm,
                                                                      double sign) {
        if (gi == groups.size()) {
          images.push_back({perm, sign});
          return;
        }
        std::vector<int> g = groups[gi], img = g;
        do {
          std::vector<int> p2 = perm;
          for (size_t k = 0; k < g.size(); ++k) p2[g[k]] = img[k];
          // parity of img relative to g
          int inv = 0;
          for (size_t x = 0; x < img.size(); ++x)
            for (size_t y = x + 1; y < img.size(); ++y)
              if (img[x] > img[y]) ++inv;
          gen(gi + 1, p2, (inv & 1) ? -sign : sign);
        } while (std::next_permutation(img.begin(), img.end()));
      };
      gen(0, {0, 1, 2, 3, 4, 5}, 1.0);
      const int64_t cols = len[0], rows = cols ? count / cols : 0;
      for (size_t ii = 0; ii < images.size(); ++ii) {
        const auto& perm = images[ii].first;
        // position p (axis values) lands at q with q[a] = p[perm[a]]:
        // offset = sum_b p[b] * str[perm^-1 b]
        int64_t pstr[6];
        for (int a = 0; a < 6; ++a) pstr[perm[a]] = str[a];
        PackPiece pc{};
        pc.rows = rows;
        pc.cols = cols;
        pc.packed_off = off;
        pc.sign = images[ii].second;
        pc.image = ii != 0;
        pc.blk = (int)pack_block.size() - 1;
        pc.t_full_row = (int64_t)host.size();
        // rows: Fortran slots 1..5, slot 1 fastest
        std::vector<int64_t> r(6, 0);
        for (int64_t m = 0; m < rows; ++m) {
          int64_t o = 0;
          for (int sl = 1; sl < 6; ++sl) o += (lo[sl] + r[sl]) * pstr[5 - sl];
          host.push_back(o);
          for (int sl = 1; sl < 6; ++sl) {
            if (++r[sl] < len[sl]) break;
            r[sl] = 0;
          }
        }
        pc.t_full_col = (int64_t)host.size();
        for (int64_t n = 0; n < cols; ++n) host.push_back((lo[0] + n) * pstr[5]);
        pc.t_pk_row = (int64_t)host.size();
        for (int64_t m = 0; m < rows; ++m) host.push_back(m * cols);
        pc.t_pk_col = (int64_t)host.size();
        for (int64_t n = 0; n < cols; ++n) host.push_back(n);
        // the full-block base offset is added at launch time
        pack_pieces.push_back(pc);
      }
      off += count;
    }
    ext_size = off;
    HIPCHK(hipMalloc(&pk_tab, sizeof(int64_t) * std::max<size_t>(host.size(), 1)));
    if (!host.empty())
      HIPCHK(hipMemcpy(pk_tab, host.data(), sizeof(int64_t) * host.size(), hipMemcpyHostToDevice));
  }
  std::vector<std::string> pack_block;   // full block of each packed piece group

  // full-storage T3 -> ccq's packed layout at dst (T1/T2 copied as they are)
  void pack(double* dst) {
    const int64_t head = blk_off.at("t3aaa");
    HIPCHK(ccqk::launch_copy(stream, dst, tvec.p, head));
    for (const PackPiece& pc : pack_pieces) {
      if (pc.image) continue;
      const double* full = tvec.p + blk_off.at(pack_block[pc.blk]);
      HIPCHK(ccqk::launch_gather64(stream, pc.rows, pc.cols, full, pk_tab + pc.t_full_row,
                                   pk_tab + pc.t_full_col, dst + pc.packed_off,
                                   pk_tab + pc.t_pk_row, pk_tab + pc.t_pk_col, 1.0));
    }
  }

  // ccq's packed layout at src -> full storage: every stored value to its
  // tuple and, with the permutation sign, to all antisymmetric images; Q-space
  // tuples are zero
  void unpack(const double* src) {
    const int64_t head = blk_off.at("t3aaa");
    HIPCHK(ccqk::launch_copy(stream, tvec.p, src, head));
    HIPCHK(ccqk::launch_fill(stream, tvec.p + head, t_size - head, 0.0));
    for (const PackPiece& pc : pack_pieces) {
      double* full = tvec.p + blk_off.at(pack_block[pc.blk]);
      HIPCHK(ccqk::launch_gather64(stream, pc.rows, pc.cols, src + pc.packed_off,
                                   pk_tab + pc.t_pk_row, pk_tab + pc.t_pk_col, full,
                                   pk_tab + pc.t_full_row, pk_tab + pc.t_full_col, pc.sign));
    }
  }

  // -------------------------------------------------------------------------
  // IR loading

  static void parse_images(std::istringstream& ls, UpdateStep& u) {
    for (int g = 0; g < u.nimg; ++g) {
      std::string tok;
      ls >> tok;
      auto colon = tok.find(':');
      if (colon == std::string::npos) throw Fail("bad image token " + tok);
      std::string ps = tok.substr(0, colon);
      double sg = std::stod(tok.substr(colon + 1));
      std::vector<int> p(8, 0);
      std::istringstream pss(ps);
      std::string x;
      int k = 0;
      while (std::getline(pss, x, ',') && k < 8) p[k++] = std::stoi(x);
      u.perms.insert(u.perms.end(), p.begin(), p.end());
      u.signs.push_back(sg);
    }
  }

  // Load one IR.  Intermediates of a second IR get the prefix ns so that the
  // programs never share scratch; views and amplitude blocks with the same
  // id are the same tensor.
  void load_ir(const char* text, const std::string& ns = "") {
    std::istringstream in(text);
    std::string line;
    Section* cur = nullptr;
    std::string prog = "sweep";
    std::map<std::string, int> local;
    auto tid = [&](const std::string& n) {
      auto it = local.find(n);
      if (it == local.end()) throw Fail("IR names undeclared tensor " + n);
      return it->second;
    };
    while (std::getline(in, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::istringstream ls(line);
      std::string kw;
      ls >> kw;
      if (kw == "ccq-ir" || kw == "end") continue;
      if (kw == "program") {
        ls >> prog;
        continue;
      }
      if (kw == "tensor") {
        Tensor t;
        std::string kind;
        ls >> t.id >> kind;
        if (kind == "inter") {
          t.kind = TKind::Inter;
          ls >> t.rank;
          for (int i = 0; i < t.rank; ++i) {
            std::string s;
            ls >> s;
            t.dims[i] = dimsym(s);
            t.spec[i] = s;
          }
        } else if (kind == "view") {
          t.kind = TKind::View;
          ls >> t.base >> t.rank;
          for (int i = 0; i < t.rank; ++i) ls >> t.spec[i];
        } else if (kind == "amp" || kind == "resid") {
          t.kind = kind == "amp" ? TKind::Amp : TKind::Resid;
          ls >> t.base;
          auto& d = blk_dims.at(dimkey(t.base));
          t.rank = (int)d.size();
          for (int i = 0; i < t.rank; ++i) t.dims[i] = d[i];
        } else {
          throw Fail("bad tensor kind " + kind);
        }
        const std::string name = t.id;
        if (t.kind == TKind::Inter) t.id = ns + t.id;
        auto hit = tindex.find(t.id);
        if (hit != tindex.end()) {
          const Tensor& o = tensors[hit->second];
          bool same = o.kind == t.kind && o.base == t.base && o.rank == t.rank &&
                      t.kind != TKind::Inter && t.kind != TKind::Resid;
          for (int i = 0; same && i < t.rank; ++i) same = o.spec[i] == t.spec[i];
          if (!same) throw Fail("IR redeclares tensor " + t.id);
          local[name] = hit->second;
        } else {
          tindex[t.id] = (int)tensors.size();
          local[name] = (int)tensors.size();
          tensors.push_back(t);
        }
      } else if (kw == "anti") {
        std::string id, pq;
        ls >> id;
        Tensor& t = tensors[local.at(id)];
        while (ls >> pq) {
          const auto c = pq.find(',');
          if (c == std::string::npos) throw Fail("bad anti record for " + id);
          const int p = std::stoi(pq.substr(0, c)), q = std::stoi(pq.substr(c + 1));
          if (p < 0 || q <= p || q >= t.rank || t.dims[p] != t.dims[q]) throw Fail("bad anti pair for " + id);
          if (std::find(t.anti.begin(), t.anti.end(), std::make_pair(p, q)) == t.anti.end())
            t.anti.emplace_back(p, q);
        }
      } else if (kw == "section") {
        sections.emplace_back();
        cur = &sections.back();
        ls >> cur->name;
        cur->prog = prog;
      } else if (kw == "op") {
        Op op;
        std::string out, a, b;
        int assign, nin;
        ls >> out >> assign >> op.coef >> nin;
        op.out = tid(out);
        op.assign = assign != 0;
        ls >> a >> op.la;
        op.a = tid(a);
        if (nin == 2) {
          ls >> b >> op.lb;
          op.b = tid(b);
          op.gemm = true;
        }
        ls >> op.lc;
        cur->kind = Section::Compute;
        cur->ops.push_back(op);
      } else if (kw == "update") {
        cur->kind = Section::Update;
        std::string rid;
        ls >> cur->upd.block >> rid >> cur->upd.nimg;
        cur->upd.resid = tid(rid);
        parse_images(ls, cur->upd);
        std::string tok;
        while (ls >> tok) {
          if (tok == "mask") {
            ls >> cur->upd.vneed >> cur->upd.vforbid;
            cur->upd.masked = true;
            if (!active) throw Fail("masked update outside an active-space program");
          } else if (tok == "left") {
            cur->upd.left = true;
          } else {
            throw Fail("bad update record tail " + tok);
          }
        }
      } else if (kw == "cct3") {
        cur->kind = Section::Cct3;
        std::string l3, m3;
        ls >> cur->upd.block >> l3 >> m3 >> cur->upd.nimg;
        cur->upd.resid = tid(l3);
        cur->upd.m3 = tid(m3);
        parse_images(ls, cur->upd);
      } else if (kw == "copy") {
        cur->kind = Section::Copy;
        ls >> cur->dst >> cur->src;
      } else {
        throw Fail("bad IR keyword " + kw);
      }
    }
  }

  // -------------------------------------------------------------------------
  // planning

  void resolve_tensors() {
    for (auto& t : tensors) {
      if (t.kind == TKind::Inter) {
        int64_t s = 1;
        for (int i = t.rank - 1; i >= 0; --i) {
          t.strides[i] = s;
          s *= t.dims[i];
        }
        inter_buf.emplace_back(new DevBuf);
        inter_buf.back()->alloc(std::max<int64_t>(s, 1));
        t.ptr = inter_buf.back()->p;
        // elements no product writes (symmetry-forbidden blocks) stay zero
        HIPCHK(hipMemset(t.ptr, 0, sizeof(double) * std::max<int64_t>(s, 1)));
      } else if (t.kind == TKind::Amp || t.kind == TKind::Resid) {
        int64_t s = 1;
        for (int i = t.rank - 1; i >= 0; --i) {
          t.strides[i] = s;
          s *= t.dims[i];
        }
        if (t.kind == TKind::Amp && t.base[0] == 'l') {
          const std::string k = dimkey(t.base);
          if (!tail || blk_off.at(k) + s > l_size) throw Fail("left amplitude block outside l_vec");
          t.ptr = lvec.p + blk_off.at(k);
        } else if (t.kind == TKind::Amp) {
          if (blk_off.at(t.base) + s > dev_t) throw Fail("amplitude block beyond device t_vec");
          t.ptr = tvec.p + blk_off.at(t.base);
        } else {
          auto& rb = resid_buf[t.base];
          if (!rb) {
            rb.reset(new DevBuf);
            rb->alloc(s);
            HIPCHK(hipMemset(rb->p, 0, sizeof(double) * std::max<int64_t>(s, 1)));
          }
          t.ptr = rb->p;
        }
      } else {
        // view into a full-range base array, materialised densely
        double* basep = nullptr;
        int64_t nb = 0;
        if (t.base == "Vaa") basep = Vaa.p, nb = 4;
        else if (t.base == "Vab") basep = Vab.p, nb = 4;
        else if (t.base == "Vbb") basep = vbb_ptr, nb = 4;
        else if (t.base == "Fa") basep = Fa.p, nb = 2;
        else if (t.base == "Fb") basep = Fb.p, nb = 2;
        else throw Fail("unknown base " + t.base);
        if (nb != t.rank) throw Fail("view rank mismatch for " + t.id);
        int64_t s = 1;
        int64_t off = 0;
        for (int i = t.rank - 1; i >= 0; --i) {
          t.src_strides[i] = s;
          const char sp = t.spec[i][0];
          const bool beta = t.spec[i][1] == 'b';
          const int64_t occ = beta ? ob : oa;
          t.dims[i] = sp == 'o' ? occ : nc - occ;
          if (sp == 'v') off += occ * s;
          s *= nc;
        }
        t.src = basep + off;
        // dense copy shared by identical views
        std::string key = std::to_string((uintptr_t)basep);
        for (int i = 0; i < t.rank; ++i) key += t.spec[i];
        auto hit = view_key.find(key);
        const std::vector<int> order = view_order(t);
        int64_t d = 1;
        for (int i = t.rank - 1; i >= 0; --i) {
          t.strides[order[i]] = d;
          d *= t.dims[order[i]];
        }
        if (hit != view_key.end()) {
          t.ptr = hit->second;
        } else {
          view_buf.emplace_back(new DevBuf);
          view_buf.back()->alloc(std::max<int64_t>(d, 1));
          t.ptr = view_buf.back()->p;
          view_key[key] = t.ptr;
          view_owner.push_back((int)(&t - &tensors[0]));
        }
      }
    }
    plan_gathers();
  }

  std::map<std::string, double*> view_key;
  std::vector<int> view_owner;

  // Storage order of a view's dense copy (outermost index first).  A view
  // whose every GEMM consumer contracts it with a small operand (at most
  // 1024 elements: the T1-dressing products) keeps the indices summed by its
  // first such consumer outermost, so that the rows those products read are
  // contiguous; any other view keeps its index order.  A function of the
  // view's base and spec, so identical views sharing a copy agree.
  // CCQ_HIP_VIEWPERM=0 keeps every view in index order.
  bool view_perm = true;

  std::vector<int> view_order(const Tensor& t) const {
    std::vector<int> ident(t.rank);
    for (int i = 0; i < t.rank; ++i) ident[i] = i;
    if (!view_perm) return ident;
    auto same_view = [&](int x) {
      const Tensor& u = tensors[x];
      if (u.kind != TKind::View || u.base != t.base || u.rank != t.rank) return false;
      for (int i = 0; i < t.rank; ++i)
        if (u.spec[i] != t.spec[i]) return false;
      return true;
    };
    std::string labels, lc;
    bool any = false;
    for (const auto& sec : sections)
      for (const auto& op : sec.ops) {
        if (!op.gemm) continue;
        const bool va = same_view(op.a), vb = same_view(op.b);
        if (!va && !vb) continue;
        if (va && vb) return ident;
        const Tensor& other = tensors[va ? op.b : op.a];
        if (other.kind == TKind::View) return ident;
        int64_t n = 1;
        for (int i = 0; i < other.rank; ++i) n *= other.dims[i];
        if (n > 1024) return ident;
        if (!any) {
          labels = va ? op.la : op.lb;
          lc = op.lc;
          any = true;
        }
      }
    if (!any) return ident;
    std::vector<int> order;
    for (int i = 0; i < t.rank; ++i)
      if (lc.find(labels[i]) == std::string::npos) order.push_back(i);
    for (int i = 0; i < t.rank; ++i)
      if (lc.find(labels[i]) != std::string::npos) order.push_back(i);
    return order;
  }

  // tables that copy each distinct view out of its base array
  void plan_gathers() {
    std::vector<int64_t> host;
    for (int ti : view_owner) {
      const Tensor& t = tensors[ti];
      std::vector<int64_t> ext(t.dims, t.dims + t.rank);
      Gather g{};
      g.t = ti;
      g.cols = ext.back();
      g.rows = 1;
      for (int i = 0; i + 1 < t.rank; ++i) g.rows *= ext[i];
      std::vector<int64_t> sr((size_t)g.rows), dr((size_t)g.rows), sc((size_t)g.cols), dc((size_t)g.cols);
      std::vector<int64_t> idx(t.rank > 1 ? t.rank - 1 : 1, 0);
      for (int64_t r = 0; r < g.rows; ++r) {
        int64_t so = 0, dof = 0;
        for (int i = 0; i + 1 < t.rank; ++i) so += idx[i] * t.src_strides[i], dof += idx[i] * t.strides[i];
        sr[r] = so;
        dr[r] = dof;
        for (int i = t.rank - 2; i >= 0; --i) {
          if (++idx[i] < ext[i]) break;
          idx[i] = 0;
        }
      }
      for (int64_t c = 0; c < g.cols; ++c) sc[c] = c * t.src_strides[t.rank - 1], dc[c] = c * t.strides[t.rank - 1];
      g.t_src_row = (int64_t)host.size(); host.insert(host.end(), sr.begin(), sr.end());
      g.t_dst_row = (int64_t)host.size(); host.insert(host.end(), dr.begin(), dr.end());
      g.t_src_col = (int64_t)host.size(); host.insert(host.end(), sc.begin(), sc.end());
      g.t_dst_col = (int64_t)host.size(); host.insert(host.end(), dc.begin(), dc.end());
      gathers.push_back(g);
    }
    HIPCHK(hipMalloc(&gather_tab, sizeof(int64_t) * std::max<size_t>(host.size(), 1)));
    if (!host.empty())
      HIPCHK(hipMemcpy(gather_tab, host.data(), sizeof(int64_t) * host.size(), hipMemcpyHostToDevice));
  }

  void gather_views() {
    for (auto& g : gathers) {
      const Tensor& t = tensors[g.t];
      HIPCHK(ccqk::launch_gather64(stream, g.rows, g.cols, t.src, gather_tab + g.t_src_row,
                                   gather_tab + g.t_src_col, t.ptr, gather_tab + g.t_dst_row,
                                   gather_tab + g.t_dst_col));
    }
    HIPCHK(hipStreamSynchronize(stream));
    views32_stale = true;
  }

  // Float shadows of the tensors the sweep reads or writes (intermediates
  // and residuals zeroed, so elements no product writes stay zero) and
  // the cast of the views; outside any graph capture.
  void prep_f32() {
    if (!f32_ready) {
      auto alloc32 = [&](int64_t n) {
        float* p = nullptr;
        const size_t bytes = sizeof(float) * (size_t)std::max<int64_t>(n, 1);
        HIPCHK(hipMalloc(&p, bytes));
        f32_bufs.push_back(p);
        HIPCHK(hipMemset(p, 0, bytes));
        return p;
      };
      tvec32 = alloc32(dev_t);
      std::set<int> used;
      for (const auto& sec : sections) {
        if (sec.prog != "sweep") continue;
        if (sec.kind == Section::Update) used.insert(sec.upd.resid);
        for (const auto& op : sec.ops) {
          for (int t : {op.out, op.a, op.b, op.fuse_a})
            if (t >= 0) used.insert(t);
          // an op into the amplitudes (the closed-shell permuted copies) runs
          // on both storages, so it may read only the amplitudes
          if (tensors[op.out].kind == TKind::Amp &&
              (op.gemm || tensors[op.a].kind != TKind::Amp || op.fuse_a >= 0))
            throw Fail("sweep op into " + tensors[op.out].id + " has no FP32-storage form");
        }
      }
      std::map<double*, float*> shadow;
      for (int ti : used) {
        Tensor& t = tensors[ti];
        if (t.kind == TKind::Amp) {
          if (t.base[0] == 'l') throw Fail("left amplitudes in the sweep");
          t.p32 = tvec32 + blk_off.at(t.base);
          continue;
        }
        auto& s = shadow[t.ptr];
        if (!s) {
          int64_t n = 1;
          for (int i = 0; i < t.rank; ++i) n *= t.dims[i];
          s = alloc32(n);
        }
        t.p32 = s;
      }
      f32_ready = true;
      views32_stale = true;
    }
    if (views32_stale) {
      std::set<float*> done;
      for (const auto& t : tensors) {
        if (t.kind != TKind::View || !t.p32 || !done.insert(t.p32).second) continue;
        int64_t n = 1;
        for (int i = 0; i < t.rank; ++i) n *= t.dims[i];
        HIPCHK(ccqk::launch_cast_f32(stream, t.p32, t.ptr, n));
      }
      HIPCHK(hipStreamSynchronize(stream));
      views32_stale = false;
    }
  }

  // storage of tensor t in the program being run
  double* sp(int t) const {
    if (run_prec != 2) return tensors[t].ptr;
    if (!tensors[t].p32) throw Fail("tensor " + tensors[t].id + " has no float shadow");
    return reinterpret_cast<double*>(tensors[t].p32);
  }

  // Offsets of the index combinations of `subset` (row-major in subset
  // order).  With a pair (px < py, positions in subset), only combinations
  // with idx[px] < idx[py] are listed; swap gives the offsets of the mirrored
  // element (the two indices exchanged).
  static std::vector<int64_t> offsets(const Tensor& t, const std::string& labels,
                                      const std::string& subset, int px = -1, int py = -1,
                                      bool swap = false) {
    std::vector<int64_t> ext, str;
    for (char c : subset) {
      auto pos = labels.find(c);
      if (pos == std::string::npos) throw Fail("label not in tensor");
      ext.push_back(t.dims[pos]);
      str.push_back(t.strides[pos]);
    }
    if (swap) std::swap(str[px], str[py]);
    int64_t total = 1;
    for (auto e : ext) total *= e;
    std::vector<int64_t> out;
    out.reserve((size_t)std::max<int64_t>(total, 1));
    if (subset.empty()) return std::vector<int64_t>(1, 0);
    std::vector<int64_t> idx(ext.size(), 0);
    for (int64_t n = 0; n < total; ++n) {
      if (px < 0 || idx[px] < idx[py]) {
        int64_t o = 0;
        for (size_t k = 0; k < ext.size(); ++k) o += idx[k] * str[k];
        out.push_back(o);
      }
      for (int k = (int)ext.size() - 1; k >= 0; --k) {
        if (++idx[k] < ext[k]) break;
        idx[k] = 0;
      }
    }
    if (out.empty()) out.push_back(0);
    return out;
  }

  // (virtual?, spin) of position pos of tensor t
  void pos_space(const Tensor& t, int pos, bool& vir, int& spin) const {
    if (t.kind == TKind::View || t.kind == TKind::Inter) {
      vir = t.spec[pos][0] == 'v';
      spin = t.spec[pos][1] == 'b' ? 1 : 0;
      return;
    }
    const int half = t.rank / 2;
    const std::string spins = dimkey(t.base).substr(2);
    vir = pos >= half;
    spin = spins[pos % half] == 'b' ? 1 : 0;
  }

  // Irrep (XOR of the orbital irreps) of every entry of
  // offsets(t, labels, subset, px, py), same enumeration
  std::vector<int> irreps_of(const Tensor& t, const std::string& labels, const std::string& subset,
                             int px = -1, int py = -1) const {
    std::vector<int64_t> ext, base;
    for (char c : subset) {
      const auto pos = labels.find(c);
      if (pos == std::string::npos) throw Fail("label not in tensor");
      bool vir;
      int spin;
      pos_space(t, (int)pos, vir, spin);
      ext.push_back(t.dims[pos]);
      base.push_back(vir ? (spin ? ob : oa) : 0);
    }
    if (subset.empty()) return std::vector<int>(1, 0);
    int64_t total = 1;
    for (auto e : ext) total *= e;
    std::vector<int> out;
    std::vector<int64_t> idx(ext.size(), 0);
    for (int64_t n = 0; n < total; ++n) {
      if (px < 0 || idx[px] < idx[py]) {
        int g = 0;
        for (size_t k = 0; k < ext.size(); ++k) g ^= orb_irrep[(size_t)(base[k] + idx[k])];
        out.push_back(g);
      }
      for (int k = (int)ext.size() - 1; k >= 0; --k) {
        if (++idx[k] < ext[k]) break;
        idx[k] = 0;
      }
    }
    if (out.empty()) out.push_back(0);
    return out;
  }

  // Active-space class of every entry of offsets(t, labels, subset, px, py)
  // (same enumeration) for a triples residual block t: bit 0 an active
  // occupied index, bit 1 an active alpha virtual, bit 2 an active beta
  // virtual among the subset's indices.
  std::vector<int> pclasses(const Tensor& t, const std::string& labels, const std::string& subset,
                            int px = -1, int py = -1) const {
    std::vector<int64_t> ext;
    std::vector<int> kind;   // 0 occupied, 1 alpha virtual, 2 beta virtual
    const int half = t.rank / 2;
    const std::string spins = dimkey(t.base).substr(2);
    for (char c : subset) {
      const auto pos = labels.find(c);
      if (pos == std::string::npos) throw Fail("label not in tensor");
      ext.push_back(t.dims[pos]);
      kind.push_back((int)pos < half ? 0 : (spins[pos % half] == 'b' ? 2 : 1));
    }
    if (subset.empty()) return std::vector<int>(1, 0);
    int64_t total = 1;
    for (auto e : ext) total *= e;
    std::vector<int> out;
    std::vector<int64_t> idx(ext.size(), 0);
    for (int64_t n = 0; n < total; ++n) {
      if (px < 0 || idx[px] < idx[py]) {
        int cls = 0;
        for (size_t k = 0; k < ext.size(); ++k) {
          if (kind[k] == 0) cls |= idx[k] >= occ_start ? 1 : 0;
          else if (idx[k] < vir_end[kind[k] - 1]) cls |= kind[k] == 1 ? 2 : 4;
        }
        out.push_back(cls);
      }
      for (int k = (int)ext.size() - 1; k >= 0; --k) {
        if (++idx[k] < ext[k]) break;
        idx[k] = 0;
      }
    }
    if (out.empty()) out.push_back(0);
    return out;
  }

  // The output of a triples residual product restricted to the tuples a
  // masked update reads (an active occupied index, an active virtual of a
  // spin in vneed, none of a spin in vforbid; jacobi_update's test): row and
  // column index lists of the rectangles that cover exactly those elements
  // (rows grouped by the set of column classes they need).
  static std::vector<std::pair<std::vector<int64_t>, std::vector<int64_t>>> pspace_parts(
      const std::vector<int>& rcl, const std::vector<int>& ccl, int vneed, int vforbid) {
    auto pred = [&](int c) {
      const bool va = c & 2, vb = c & 4;
      const bool need = ((vneed & 1) && va) || ((vneed & 2) && vb);
      const bool forbid = ((vforbid & 1) && va) || ((vforbid & 2) && vb);
      return (c & 1) && need && !forbid;
    };
    std::map<int, std::vector<int64_t>> rows_by_set;
    for (size_t m = 0; m < rcl.size(); ++m) {
      int set = 0;
      for (int c = 0; c < 8; ++c)
        if (pred(rcl[m] | c)) set |= 1 << c;
      if (set) rows_by_set[set].push_back((int64_t)m);
    }
    std::vector<std::pair<std::vector<int64_t>, std::vector<int64_t>>> parts;
    for (auto& kv : rows_by_set) {
      std::vector<int64_t> cols;
      for (size_t n = 0; n < ccl.size(); ++n)
        if ((kv.first >> ccl[n]) & 1) cols.push_back((int64_t)n);
      if (!cols.empty()) parts.emplace_back(kv.second, cols);
    }
    return parts;
  }

  // A piece of a product: its rows, columns and summed entries (empty ks:
  // all of them)
  struct Part {
    std::vector<int64_t> rows, cols, ks;
    bool no_k = false;   // a block no summed entry reaches (an assigning product writes zeros)
  };

  static std::vector<int64_t> iota64(int64_t n) {
    std::vector<int64_t> v((size_t)std::max<int64_t>(n, 0));
    for (int64_t i = 0; i < n; ++i) v[(size_t)i] = i;
    return v;
  }

  // Irrep blocks of each part: for every irrep g the rows, columns and
  // summed entries of irrep g (a totally symmetric A[m,k] vanishes unless
  // irrep(m) = irrep(k), and B[k,n] unless irrep(k) = irrep(n)); blocks
  // without rows or columns are dropped, and so are blocks of a product
  // without summed entries unless the product assigns its output (then the
  // block is kept, no_k, and written with zeros)
  static std::vector<Part> irrep_split(const std::vector<Part>& parts, const std::vector<int>& rir,
                                       const std::vector<int>& cir, const std::vector<int>* kir,
                                       bool assign) {
    std::vector<Part> out;
    for (const Part& p : parts) {
      for (int g = 0; g < 8; ++g) {
        Part q;
        for (auto m : p.rows)
          if (rir[(size_t)m] == g) q.rows.push_back(m);
        for (auto n : p.cols)
          if (cir[(size_t)n] == g) q.cols.push_back(n);
        if (kir) {
          for (size_t k = 0; k < kir->size(); ++k)
            if ((*kir)[k] == g) q.ks.push_back((int64_t)k);
          if (q.ks.empty()) {
            if (!assign) continue;
            q.no_k = true;
          }
        }
        if (!q.rows.empty() && !q.cols.empty()) out.push_back(std::move(q));
      }
    }
    return out;
  }

  static std::vector<int64_t> pick(const std::vector<int64_t>& v, const std::vector<int64_t>& sel) {
    std::vector<int64_t> out;
    out.reserve(sel.size());
    for (auto i : sel) out.push_back(v[(size_t)i]);
    return out;
  }

  static bool is_anti(const Tensor& t, const std::string& labels, char x, char y) {
    const auto p = labels.find(x), q = labels.find(y);
    if (p == std::string::npos || q == std::string::npos) return false;
    const auto pr = std::make_pair((int)std::min(p, q), (int)std::max(p, q));
    return std::find(t.anti.begin(), t.anti.end(), pr) != t.anti.end();
  }

  // first pair of labels of `sub` (extent at least 2) antisymmetric in both
  // (X, lx) and, when given, (Y, ly): positions in sub, or {-1, -1}
  static std::pair<int, int> find_pair(const std::string& sub, const Tensor& X, const std::string& lx,
                                       const Tensor* Y, const std::string& ly) {
    for (int i = 0; i < (int)sub.size(); ++i)
      for (int j = i + 1; j < (int)sub.size(); ++j)
        if (X.dims[lx.find(sub[i])] >= 2 && is_anti(X, lx, sub[i], sub[j]) &&
            (!Y || is_anti(*Y, ly, sub[i], sub[j])))
          return {i, j};
    return {-1, -1};
  }

  // CCQ_HIP_PACK=0 computes every element of antisymmetric pairs
  bool pack_pairs = true;
  static constexpr int64_t kPackMinK = 64;

  static int64_t stride_of(const Tensor& t, const std::string& labels, char c) {
    auto pos = labels.find(c);
    return t.strides[pos];
  }

  int64_t ws_need = 0;

  // k * stride for every k (the offset table of a summed index run that is
  // contiguous-by-stride in memory), or -1
  static int64_t affine_stride(const std::vector<int64_t>& t) {
    if (t.empty() || t[0] != 0) return -1;
    if (t.size() == 1) return 0;
    const int64_t s = t[1];
    for (size_t k = 0; k < t.size(); ++k)
      if (t[k] != (int64_t)k * s) return -1;
    return s;
  }

  static std::string plan_tag(const ccqk::GemmArgs& g) {
    std::string t;
    if (g.kind == 1)
      t = "r" + std::to_string(g.rm) + std::to_string(g.rn) + "w" + std::to_string(g.wk) +
          (g.trans ? "t" : "");
    else
      t = std::to_string(g.bm) + "x" + std::to_string(g.bn) + "k" + std::to_string(g.bk) +
          (g.ksplit == 2 ? "d" : "") + (g.trans ? "t" : "");
    return t + "s" + std::to_string(g.nsplit);
  }

  // Measuring autotuner.  For every product launched on its own, the
  // planner's kTuneCandidates best plans by modelled cycles are each run, in
  // two rounds, once to warm up and kTuneReps times timed with HIP events on
  // the context's stream (before the first sweep: products only write
  // intermediates and residual accumulators, which every sweep recomputes),
  // and the fastest by its better round is kept.  Results are shared by every product with the
  // same shape and operand layouts for the life of the process, so every
  // context of a process makes the same choices.
  static constexpr int kTuneCandidates = 6;
  static constexpr int kTuneReps = 3;

  struct TuneCache {
    std::mutex mu;
    std::map<std::string, ccqk::GemmArgs> plans;
  };
  static TuneCache& tune_cache() {
    static TuneCache c;
    return c;
  }

  static std::string tune_key(const Op& op) {
    const ccqk::GemmArgs& g = op.plan;
    std::ostringstream k;
    k << g.M << 'x' << g.N << 'x' << g.K << ':' << g.a_kfast << g.b_kfast << g.c_mfast << g.c_nfast << ':'
      << g.kA_aff << ':' << g.kB_aff << ':' << (op.assign ? 0 : 1) << (op.t_rowC2 >= 0) << (op.t_colC2 >= 0);
    return k.str();
  }

  static void copy_plan(ccqk::GemmArgs& dst, const ccqk::GemmArgs& src) {
    dst.kind = src.kind;
    dst.trans = src.trans;
    dst.bm = src.bm;
    dst.bn = src.bn;
    dst.bk = src.bk;
    dst.rm = src.rm;
    dst.rn = src.rn;
    dst.wk = src.wk;
    dst.nsplit = src.nsplit;
    dst.kchunk = src.kchunk;
    dst.group = src.group;
    dst.ksplit = src.ksplit;
  }

  // prec 0: plans timed in FP64 on the FP64 tensors (op.plan); prec 2: the
  // sweep's products timed in FP32 storage on the float shadows (op.plan32)
  void autotune(const std::vector<int>& host32, int prec = 0) {
    run_prec = prec;
    int* dt = nullptr;
    HIPCHK(hipMalloc(&dt, sizeof(int) * std::max<size_t>(host32.size(), 1)));
    HIPCHK(hipMemcpy(dt, host32.data(), sizeof(int) * host32.size(), hipMemcpyHostToDevice));
    DevBuf tws;
    tws.alloc(kWsCap);
    int* saved_tab = tables32;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    tables32 = dt;
    cur_ws = tws.p;
    auto restore = [&] {
      tables32 = saved_tab;
      cur_ws = nullptr;
      run_prec = 0;
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      (void)hipFree(dt);
    };
    try {
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      TuneCache& cache = tune_cache();
      for (auto& sec : sections) {
        if (sec.kind != Section::Compute || (prec == 2 && sec.prog != "sweep")) continue;
        for (auto& op : sec.ops) {
          if (!op.gemm || op.has_gplan || op.K == 0) continue;
          const std::string key = (prec == 2 ? "p2:" : "") + tune_key(op);
          if (prec == 2) {
            op.plan32 = op.plan;
            op.has_plan32 = true;
          }
          ccqk::GemmArgs& dst = prec == 2 ? op.plan32 : op.plan;
          {
            std::lock_guard<std::mutex> lk(cache.mu);
            auto hit = cache.plans.find(key);
            if (hit != cache.plans.end()) {
              copy_plan(dst, hit->second);
              continue;
            }
          }
          std::vector<std::pair<double, ccqk::GemmArgs>> ranked;
          ccqk::GemmArgs model = op.plan;
          ccqk::gemm_plan_ranked(model, kWsCap, &ranked);
          std::vector<ccqk::GemmArgs> cands;
          for (auto& rc : ranked) {
            const ccqk::GemmArgs& c = rc.second;
            bool dup = false;
            for (auto& d : cands)
              dup = dup || (d.kind == c.kind && d.trans == c.trans && d.bm == c.bm && d.bn == c.bn &&
                            d.bk == c.bk && d.rm == c.rm && d.rn == c.rn && d.wk == c.wk &&
                            d.nsplit == c.nsplit && d.ksplit == c.ksplit);
            if (!dup) cands.push_back(c);
            if ((int)cands.size() == kTuneCandidates) break;
          }
          // two rounds over the candidates, each keeping its faster round, so
          // that the order of the candidates (clock ramp, caches warmed by
          // the previous one) does not decide
          std::vector<double> t_best(cands.size(), 1e300);
          for (int round = 0; round < 2; ++round)
            for (size_t ci = 0; ci < cands.size(); ++ci) {
              Op trial = op;
              copy_plan(trial.plan, cands[ci]);
              trial.has_plan32 = false;
              ccqk::GemmArgs a = gemm_args(trial);
              HIPCHK(ccqk::launch_gemm(stream, a));
              HIPCHK(hipEventRecord(e0, stream));
              for (int r = 0; r < kTuneReps; ++r) HIPCHK(ccqk::launch_gemm(stream, a));
              HIPCHK(hipEventRecord(e1, stream));
              HIPCHK(hipEventSynchronize(e1));
              float ms = 0.0f;
              HIPCHK(hipEventElapsedTime(&ms, e0, e1));
              t_best[ci] = std::min(t_best[ci], (double)ms);
            }
          ccqk::GemmArgs best = dst;
          double best_t = 1e300;
          for (size_t ci = 0; ci < cands.size(); ++ci)
            if (t_best[ci] < best_t) {
              best_t = t_best[ci];
              best = cands[ci];
            }
          copy_plan(dst, best);
          std::lock_guard<std::mutex> lk(cache.mu);
          cache.plans.emplace(key, best);
        }
      }
    } catch (...) {
      restore();
      throw;
    }
    restore();
    // class names follow the chosen plans
    for (auto& sec : sections)
      for (auto& op : sec.ops)
        if (op.gemm)
          op.cls = "gemm:" + std::to_string(op.M) + "x" + std::to_string(op.N) + "x" + std::to_string(op.K) +
                   "/" + plan_tag(op.plan) + (op.pack_k || op.t_rowC2 >= 0 || op.t_colC2 >= 0 ? "p" : "");
  }

  // A dressing section (T1-dressed intermediates, name "dress...") and the
  // compute section right after it in its program are planned as one: their
  // products form one dependency DAG, so the later section's products that
  // do not read the dressed intermediates can share the dressing levels.
  // The merged section keeps the later section's name (the P-space rule
  // keys on residual section names); consecutive residual sections, each
  // feeding its own masked update, stay apart.  CCQ_HIP_MERGE=0 keeps the
  // program's sections.
  bool merge_sections = true;
  void merge_compute_sections() {
    std::vector<Section> out;
    for (auto& sec : sections) {
      if (!out.empty() && sec.kind == Section::Compute && out.back().kind == Section::Compute &&
          out.back().prog == sec.prog && out.back().name.rfind("dress", 0) == 0) {
        std::vector<Op> ops = std::move(out.back().ops);
        ops.insert(ops.end(), sec.ops.begin(), sec.ops.end());
        out.back() = std::move(sec);
        out.back().ops = std::move(ops);
        continue;
      }
      out.push_back(std::move(sec));
    }
    sections = std::move(out);
  }

  void plan() {
    if (merge_sections) merge_compute_sections();
    std::vector<int> host32;
    auto push32 = [&](const std::vector<int64_t>& v) {
      int64_t o = (int64_t)host32.size();
      for (auto x : v) {
        if (x < 0 || x > 0x7fffffffLL) throw Fail("tensor offset exceeds int32 offset tables");
        host32.push_back((int)x);
      }
      return o;
    };
    std::vector<int> ih;
    std::vector<double> dh;
    lsq_n = 0;
    // CCSDt: a triples residual section feeds one masked update (the next
    // update of its block); its products are computed on that update's tuples
    std::map<const Section*, std::pair<int, int>> pmask;
    if (active && pspace) {
      for (size_t i = 0; i < sections.size(); ++i) {
        const Section& sc = sections[i];
        if (sc.kind != Section::Compute || sc.name.rfind("resid:t3", 0) != 0) continue;
        const std::string blk = sc.name.substr(6);
        for (size_t j = i + 1; j < sections.size(); ++j)
          if (sections[j].kind == Section::Update && sections[j].upd.block == blk) {
            if (sections[j].upd.masked) pmask[&sc] = {sections[j].upd.vneed, sections[j].upd.vforbid};
            break;
          }
      }
    }
    for (auto& sec : sections) {
      if (sec.kind == Section::Update || sec.kind == Section::Cct3) {
        sec.upd.dev_perm_off = (int)ih.size();
        ih.insert(ih.end(), sec.upd.perms.begin(), sec.upd.perms.end());
        sec.upd.dev_sign_off = (int)dh.size();
        dh.insert(dh.end(), sec.upd.signs.begin(), sec.upd.signs.end());
        if (sec.kind == Section::Update) {
          // canonical halves: every tuple of the half in index order, kept
          // when ascending inside each antisymmetric group
          const ccqk::UpdateParams P0 = update_params(sec.upd);
          const int half = P0.rank / 2;
          for (int h = 0; h < 2; ++h) {
            const int d0 = h * half;
            const int64_t off = (int64_t)ih.size();
            int64_t count = 0;
            std::vector<int64_t> idx(half, 0);
            int64_t total = 1;
            for (int d = 0; d < half; ++d) total *= P0.dims[d0 + d];
            for (int64_t e = 0; e < total; ++e) {
              bool canon = true;
              for (int d = 0; d + 1 < half; ++d)
                if (P0.group[d0 + d] >= 0 && P0.group[d0 + d] == P0.group[d0 + d + 1] && idx[d] >= idx[d + 1])
                  canon = false;
              if (canon) {
                for (int d = 0; d < half; ++d) ih.push_back((int)idx[d]);
                ++count;
              }
              for (int d = half - 1; d >= 0; --d) {
                if (++idx[d] < P0.dims[d0 + d]) break;
                idx[d] = 0;
              }
            }
            if (off > 0x7fffffffLL || (int64_t)ih.size() > 0x7fffffffLL) throw Fail("update table too large");
            (h == 0 ? sec.upd.t_coc : sec.upd.t_cvc) = (int)off;
            (h == 0 ? sec.upd.n_oc : sec.upd.n_vc) = count;
          }
        }
        if (sec.kind == Section::Update && sec.upd.left) {
          sec.upd.sq_off = (int)lsq_n;
          sec.upd.sq_n = ccqk::update_blocks(update_params(sec.upd));
          lsq_n += sec.upd.sq_n;
        }
        continue;
      }
      if (sec.kind == Section::Copy) continue;
      const auto pm_it = pmask.find(&sec);
      std::vector<Op> planned;
      for (Op& op : sec.ops) {
        const Tensor& C = tensors[op.out];
        const bool restrict_out = pm_it != pmask.end() && C.kind == TKind::Resid && C.rank == 6;
        const Tensor& A = tensors[op.a];
        for (size_t i = 0; i < op.lc.size(); ++i) {
          // extents must agree label by label
          char c = op.lc[i];
          auto pa = op.la.find(c);
          if (pa != std::string::npos && A.dims[pa] != C.dims[i]) throw Fail("extent mismatch A/C");
          if (op.gemm) {
            auto pb = op.lb.find(c);
            if (pb != std::string::npos && tensors[op.b].dims[pb] != C.dims[i])
              throw Fail("extent mismatch B/C");
          }
        }
        std::string ms, ns, ks;
        if (op.gemm) {
          const Tensor& B = tensors[op.b];
          for (char c : op.la) {
            bool inC = op.lc.find(c) != std::string::npos;
            bool inB = op.lb.find(c) != std::string::npos;
            if (inC && inB) throw Fail("batch label unsupported in " + sec.name);
            if (!inC && !inB) throw Fail("partial trace unsupported in " + sec.name);
            if (!inC) ks += c;
          }
          for (char c : op.lb) {
            bool inC = op.lc.find(c) != std::string::npos;
            bool inA = op.la.find(c) != std::string::npos;
            if (!inC && !inA) throw Fail("partial trace unsupported in " + sec.name);
          }
          for (char c : op.lc) {
            if (op.la.find(c) != std::string::npos) ms += c;
            else ns += c;
          }
          for (char c : ks) {
            auto pa = op.la.find(c), pb = op.lb.find(c);
            if (A.dims[pa] != B.dims[pb]) throw Fail("extent mismatch on summed label");
          }
          // antisymmetric pairs: a summed pair antisymmetric in A and B
          // contributes twice its canonical half; an output pair carried
          // by A (or B) makes the contribution antisymmetric in it, so only
          // its canonical half is computed and the mirrored elements are
          // written with the opposite sign (accumulating ops only: the
          // diagonal is never written)
          // (output pairs only for products with K >= kPackMinK: with a short
          // K the product is store bound and the mirrored stores cost more
          // than the halved arithmetic saves)
          std::pair<int, int> pk{-1, -1}, pm{-1, -1}, pn{-1, -1};
          if (pack_pairs) {
            pk = find_pair(ks, A, op.la, &B, op.lb);
            int64_t kfull = 1;
            for (char c : ks) kfull *= A.dims[op.la.find(c)];
            if (!op.assign && kfull >= kPackMinK) {
              pm = find_pair(ms, A, op.la, nullptr, "");
              pn = find_pair(ns, B, op.lb, nullptr, "");
            }
          }
          const auto rA_all = offsets(A, op.la, ms, pm.first, pm.second);
          const auto kA = offsets(A, op.la, ks, pk.first, pk.second);
          const auto kB = offsets(B, op.lb, ks, pk.first, pk.second);
          const auto cB_all = offsets(B, op.lb, ns, pn.first, pn.second);
          const auto rC_all = offsets(C, op.lc, ms, pm.first, pm.second);
          const auto cC_all = offsets(C, op.lc, ns, pn.first, pn.second);
          std::vector<int64_t> rC2_all, cC2_all;
          if (pm.first >= 0) rC2_all = offsets(C, op.lc, ms, pm.first, pm.second, true);
          if (pn.first >= 0) cC2_all = offsets(C, op.lc, ns, pn.first, pn.second, true);
          op.pack_k = pk.first >= 0;
          if (pk.first >= 0) op.coef *= 2.0;
          std::vector<Part> parts(1);
          parts[0].rows = iota64((int64_t)rA_all.size());
          parts[0].cols = iota64((int64_t)cB_all.size());
          if (restrict_out) {
            parts.clear();
            for (auto& rc : pspace_parts(pclasses(C, op.lc, ms, pm.first, pm.second),
                                         pclasses(C, op.lc, ns, pn.first, pn.second),
                                         pm_it->second.first, pm_it->second.second))
              parts.push_back(Part{rc.first, rc.second, {}});
          }
          if (use_sym) {
            const auto kir = irreps_of(A, op.la, ks, pk.first, pk.second);
            parts = irrep_split(parts, irreps_of(C, op.lc, ms, pm.first, pm.second),
                                irreps_of(C, op.lc, ns, pn.first, pn.second), &kir, op.assign);
          }
          const bool sub = restrict_out || use_sym;
          const int pid = sub && parts.size() > 1 ? part_count++ : -1;
          const Op base = op;
          const size_t first_part = planned.size();
          for (auto& part : parts) {
          Op op = base;
          op.restricted = sub;
          op.part_of = pid;
          const auto rA = sub ? pick(rA_all, part.rows) : rA_all;
          const auto cB = sub ? pick(cB_all, part.cols) : cB_all;
          const auto rC = sub ? pick(rC_all, part.rows) : rC_all;
          const auto cC = sub ? pick(cC_all, part.cols) : cC_all;
          const auto kAp = part.no_k ? std::vector<int64_t>() : part.ks.empty() ? kA : pick(kA, part.ks);
          const auto kBp = part.no_k ? std::vector<int64_t>() : part.ks.empty() ? kB : pick(kB, part.ks);
          op.M = (int64_t)rA.size();
          op.N = (int64_t)cB.size();
          op.K = (int64_t)kAp.size();
          op.t_rowA = push32(rA);
          op.t_kA = push32(kAp);
          op.t_kB = push32(kBp);
          op.t_colB = push32(cB);
          op.t_rowC = push32(rC);
          op.t_colC = push32(cC);
          if (pm.first >= 0) op.t_rowC2 = push32(sub ? pick(rC2_all, part.rows) : rC2_all);
          if (pn.first >= 0) op.t_colC2 = push32(sub ? pick(cC2_all, part.cols) : cC2_all);
          ccqk::GemmArgs& g = op.plan;
          g.M = op.M;
          g.N = op.N;
          g.K = op.K;
          g.a_kfast = ks.empty() ? 1 : (stride_of(A, op.la, ks.back()) == 1 ? 1
                      : (!ms.empty() && stride_of(A, op.la, ms.back()) == 1 ? 0 : 1));
          g.b_kfast = ks.empty() ? 1 : (stride_of(B, op.lb, ks.back()) == 1 ? 1
                      : (!ns.empty() && stride_of(B, op.lb, ns.back()) == 1 ? 0 : 1));
          g.kA_aff = affine_stride(kAp);
          g.kB_aff = affine_stride(kBp);
          // C contiguous along m (n): its fastest m (n) label has stride 1,
          // so consecutive rows (columns) are adjacent in runs of that extent
          g.c_mfast = !ms.empty() && stride_of(C, op.lc, ms.back()) == 1 ? 1 : 0;
          g.c_nfast = !ns.empty() && stride_of(C, op.lc, ns.back()) == 1 ? 1 : 0;
          if (op.K == 0) {
            // an irrep block no summed entry reaches: the tile kernel with no
            // K slab stores zeros (the product assigns its output)
            g.kind = 0;
            g.trans = 0;
            g.bm = g.bn = 32;
            g.bk = 32;
            g.ksplit = 1;
            g.nsplit = 1;
            g.kchunk = 0;
          } else {
            ccqk::gemm_plan(g, kWsCap);
          }
          ws_need = std::max<int64_t>(ws_need, ccqk::gemm_ws_need(g));
          op.cls = "gemm:" + std::to_string(op.M) + "x" + std::to_string(op.N) + "x" +
                   std::to_string(op.K) + "/" + plan_tag(g) +
                   (op.pack_k || op.t_rowC2 >= 0 || op.t_colC2 >= 0 ? "p" : "");
          op.flops = 2.0 * (double)op.M * (double)op.N * (double)(ks.empty() ? 1 : op.K);
          op.bytes = 8.0 * ((double)op.M * op.K + (double)op.K * op.N + 2.0 * op.M * op.N);
          planned.push_back(op);
          }
          if (pid >= 0) {
            // the parts run as grouped launches: one tile shape for all
            int64_t mmax = 0, nmax = 0;
            for (size_t q = first_part; q < planned.size(); ++q) {
              mmax = std::max(mmax, planned[q].M);
              nmax = std::max(nmax, planned[q].N);
            }
            for (size_t q = first_part; q < planned.size(); ++q) {
              // a long summed range stays with its own plan (split-K, or the
              // register-direct kernel): a grouped member runs its whole K
              if (planned[q].K > kPartGroupMaxK) continue;
              ccqk::GemmArgs gp = planned[q].plan;
              gp.kind = 0;
              gp.trans = 0;
              gp.ksplit = 1;
              gp.bm = mmax >= 96 ? 64 : 32;
              gp.bn = nmax >= 96 ? 64 : 32;
              gp.bk = 32;
              gp.nsplit = 1;
              gp.kchunk = gp.K;
              planned[q].gplan = gp;
              planned[q].has_gplan = true;
            }
          }
        } else {
          if (op.la.size() != op.lc.size()) throw Fail("add op must permute its input");
          ms = op.lc.substr(0, op.lc.size() - 1);
          ns = op.lc.substr(op.lc.size() - 1);
          const auto rA_all = offsets(A, op.la, ms), cA_all = offsets(A, op.la, ns);
          const auto rC_all = offsets(C, op.lc, ms), cC_all = offsets(C, op.lc, ns);
          std::vector<Part> parts(1);
          parts[0].rows = iota64((int64_t)rA_all.size());
          parts[0].cols = iota64((int64_t)cA_all.size());
          if (restrict_out) {
            parts.clear();
            for (auto& rc : pspace_parts(pclasses(C, op.lc, ms), pclasses(C, op.lc, ns),
                                         pm_it->second.first, pm_it->second.second))
              parts.push_back(Part{rc.first, rc.second, {}});
          }
          if (use_sym)
            parts = irrep_split(parts, irreps_of(C, op.lc, ms), irreps_of(C, op.lc, ns), nullptr, op.assign);
          const bool sub = restrict_out || use_sym;
          const int pid = sub && parts.size() > 1 ? part_count++ : -1;
          const Op base = op;
          for (auto& part : parts) {
          Op op = base;
          op.restricted = sub;
          op.part_of = pid;
          const auto rA = sub ? pick(rA_all, part.rows) : rA_all;
          const auto cA = sub ? pick(cA_all, part.cols) : cA_all;
          const auto rC = sub ? pick(rC_all, part.rows) : rC_all;
          const auto cC = sub ? pick(cC_all, part.cols) : cC_all;
          op.M = (int64_t)rA.size();
          op.N = (int64_t)cA.size();
          // same layout on both sides: a flat (vectorised) axpy
          op.flat = rA == rC && cA == cC && affine_stride(cA) == 1 && affine_stride(rA) == op.N;
          op.t_rowA = push32(rA);
          op.t_kA = push32(cA);
          op.t_rowC = push32(rC);
          op.t_colC = push32(cC);
          op.flops = (double)op.M * op.N;
          op.bytes = 8.0 * 3.0 * op.M * op.N;
          op.cls = op.flat ? "add:flat" : "add:perm";
          planned.push_back(op);
          }
        }
      }
      sec.ops = std::move(planned);
    }
    if (tune) {
      autotune(host32);
      if (tune32) {
        prep_f32();
        autotune(host32, 2);
      }
      ws_need = 0;
      for (auto& sec : sections)
        for (auto& op : sec.ops)
          if (op.gemm) {
            ws_need = std::max<int64_t>(ws_need, ccqk::gemm_ws_need(op.plan));
            if (op.has_plan32) ws_need = std::max<int64_t>(ws_need, ccqk::gemm_ws_need(op.plan32));
          }
    }
    fuse_adds(push32);
    // groups first; products left alone are chained where they can be
    group_gemms(false);
    chain_gemms(host32);
    for (auto& sec : sections) {
      sec.launches.clear();
      sec.launch_cls.clear();
      sec.launch_level.clear();
    }
    group_gemms(dump_schedule);
    ws.alloc(std::max<int64_t>(ws_need, 1));
    cur_ws = ws.p;
    if (nstreams > 1) {
      HIPCHK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
      for (int k = 0; k < nstreams - 1; ++k) {
        HIPCHK(hipStreamCreateWithFlags(&side[k], hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&join_ev[k], hipEventDisableTiming));
        ws_side[k].alloc(std::max<int64_t>(ws_need, 1));
      }
    }
    if (lsq_n) lsq.alloc(lsq_n);
    HIPCHK(hipMalloc(&tables32, sizeof(int) * std::max<size_t>(host32.size(), 1)));
    if (!host32.empty())
      HIPCHK(hipMemcpy(tables32, host32.data(), sizeof(int) * host32.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&itab, sizeof(int) * std::max<size_t>(ih.size(), 1)));
    if (!ih.empty()) HIPCHK(hipMemcpy(itab, ih.data(), sizeof(int) * ih.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&dtab, sizeof(double) * std::max<size_t>(dh.size(), 1)));
    if (!dh.empty())
      HIPCHK(hipMemcpy(dtab, dh.data(), sizeof(double) * dh.size(), hipMemcpyHostToDevice));
  }

  // Fold an add op (C (+)= c A) into the GEMM that next updates C, when that
  // GEMM accumulates into C, does not read C, and A is not written in
  // between: the GEMM's epilogue adds c A (C = c A + alpha A'B' for an
  // assigning add).  Saves the add's launch and its pass over C.
  // CCQ_HIP_FUSE=0 keeps the adds.
  bool fuse = true;

  template <typename Push>
  void fuse_adds(Push&& push32) {
    if (!fuse) return;
    for (auto& sec : sections) {
      if (sec.kind != Section::Compute) continue;
      auto& ops = sec.ops;
      for (size_t i = 0; i < ops.size(); ++i) {
        Op& ad = ops[i];
        if (ad.gemm || ad.fused || ad.a == ad.out || ad.restricted) continue;
        size_t j = i + 1;
        for (; j < ops.size(); ++j) {
          const Op& q = ops[j];
          if (q.out == ad.out || q.a == ad.out || (q.gemm && q.b == ad.out)) break;
        }
        if (j == ops.size()) continue;
        Op& gm = ops[j];
        if (!gm.gemm || gm.restricted || gm.out != ad.out || gm.assign || gm.fuse_a >= 0 || gm.a == ad.out ||
            gm.b == ad.out || gm.t_rowC2 >= 0 || gm.t_colC2 >= 0)
          continue;
        bool clobbered = false;
        for (size_t k = i + 1; k < j; ++k)
          if (ops[k].out == ad.a) clobbered = true;
        if (clobbered) continue;
        // offsets of A for the GEMM's row / column label sets of C: C position
        // p carries add label lc[p], read from A at la.find(lc[p])
        const Tensor& A = tensors[ad.a];
        const Tensor& C = tensors[gm.out];
        auto table = [&](bool rows) {
          std::vector<int64_t> ext, str;
          for (size_t p = 0; p < gm.lc.size(); ++p) {
            const char c = gm.lc[p];
            const bool in_a = gm.la.find(c) != std::string::npos;
            if (in_a != rows) continue;
            const auto q = ad.la.find(ad.lc[p]);
            if (q == std::string::npos) throw Fail("add op label mismatch");
            ext.push_back(C.dims[p]);
            str.push_back(A.strides[q]);
          }
          int64_t total = 1;
          for (auto e : ext) total *= e;
          std::vector<int64_t> out((size_t)std::max<int64_t>(total, 1), 0);
          std::vector<int64_t> idx(ext.size(), 0);
          for (int64_t t = 0; t < total; ++t) {
            int64_t o = 0;
            for (size_t k = 0; k < ext.size(); ++k) o += idx[k] * str[k];
            out[t] = o;
            for (int k = (int)ext.size() - 1; k >= 0; --k) {
              if (++idx[k] < ext[k]) break;
              idx[k] = 0;
            }
          }
          return out;
        };
        auto rowD = table(true), colD = table(false);
        if ((int64_t)rowD.size() != gm.M || (int64_t)colD.size() != gm.N) continue;
        gm.t_rowD = push32(rowD);
        gm.t_colD = push32(colD);
        gm.fuse_a = ad.a;
        gm.dcoef = ad.coef;
        gm.assign = ad.assign;
        ad.fused = true;
      }
    }
  }

  // GEMM chains.  Consecutive GEMMs that accumulate into one output with the
  // same tile mapping (equal row / column offset tables, or equal after
  // transposing one product) and that the planner would split along K run as
  // one product over their concatenated K ranges (gemm_tab_chain): one
  // launch, one split-K reduction, one pass over the output.  The first
  // member may assign and carry a folded add; the others accumulate.
  // CCQ_HIP_CHAIN=0 keeps one launch per GEMM.
  bool chain = true;

  void chain_gemms(const std::vector<int>& host32) {
    if (!chain) return;
    auto same = [&](int64_t a, int64_t b, int64_t n) {
      return std::equal(host32.begin() + a, host32.begin() + a + n, host32.begin() + b);
    };
    for (auto& sec : sections) {
      if (sec.kind != Section::Compute) continue;
      auto& ops = sec.ops;
      // products that run in a grouped launch stay there
      std::set<int> grouped;
      for (const auto& L : sec.launches)
        if (L.size() > 1) grouped.insert(L.begin(), L.end());
      auto eligible = [&](int i) {
        const Op& op = ops[i];
        return op.gemm && !op.fused && !grouped.count(i) && op.plan.kind == 0 && op.t_rowC2 < 0 &&
               op.t_colC2 < 0 && op.a != op.out && op.b != op.out;
      };
      std::vector<int> live;
      for (int i = 0; i < (int)ops.size(); ++i)
        if (!ops[i].fused) live.push_back(i);
      for (size_t p = 0; p < live.size();) {
        const Op& h = ops[live[p]];
        if (!eligible(live[p]) || h.plan.nsplit < 2) {
          ++p;
          continue;
        }
        std::vector<int> mem{live[p]}, flip{0};
        int64_t ktot = h.K;
        size_t q = p + 1;
        for (; q < live.size() && mem.size() < (size_t)ccqk::kMaxChain; ++q) {
          const Op& x = ops[live[q]];
          if (!eligible(live[q]) || x.out != h.out || x.assign || x.fuse_a >= 0) break;
          // same tile mapping and, in that orientation, the same operand layouts
          int f = -1;
          if (x.M == h.M && x.N == h.N && same(x.t_rowC, h.t_rowC, h.M) && same(x.t_colC, h.t_colC, h.N) &&
              x.plan.a_kfast == h.plan.a_kfast && x.plan.b_kfast == h.plan.b_kfast)
            f = 0;
          else if (x.M == h.N && x.N == h.M && same(x.t_rowC, h.t_colC, h.N) && same(x.t_colC, h.t_rowC, h.M) &&
                   x.plan.b_kfast == h.plan.a_kfast && x.plan.a_kfast == h.plan.b_kfast)
            f = 1;
          if (f < 0) break;
          mem.push_back(live[q]);
          flip.push_back(f);
          ktot += x.K;
        }
        if (mem.size() < 2) {
          ++p;
          continue;
        }
        ccqk::GemmArgs cp = h.plan;
        cp.K = ktot;
        cp.kA_aff = cp.kB_aff = -1;
        ccqk::gemm_plan(cp, kWsCap);
        // (the chain kernel spills registers with tiles above 64 x 64)
        if (cp.kind != 0 || cp.bm * cp.bn > 64 * 64) {
          ++p;
          continue;
        }
        cp.bk = 32;
        cp.ksplit = 1;
        ws_need = std::max<int64_t>(ws_need, ccqk::gemm_ws_need(cp));
        const int c = (int)sec.chains.size();
        sec.chains.push_back(mem);
        sec.chain_flip.push_back(flip);
        sec.chain_plan.push_back(cp);
        for (int i : mem) ops[i].chain = c;
        ops[mem[0]].chain_head = true;
        p = q;
      }
    }
  }

  std::string chain_cls(const Section& sec, int c) const {
    const auto& mem = sec.chains[c];
    const ccqk::GemmArgs& cp = sec.chain_plan[c];
    const Op& h = sec.ops[mem[0]];
    std::string t = "gemm:" + std::to_string(mem.size()) + "*" + std::to_string(h.M) + "x" +
                    std::to_string(h.N) + "x" + std::to_string(cp.K / (int64_t)mem.size());
    return t + "/" + std::to_string(cp.bm) + "x" + std::to_string(cp.bn) + "k32s" +
           std::to_string(cp.nsplit) + (cp.trans ? "t" : "") + "c";
  }

  // entries of a chain, oriented like its first member run with trans
  // (1: the whole chain transposed), tile bm x bn, BK 32
  void chain_entries(const Section& sec, int c, int trans, int bm, int bn, ccqk::GemmArgs* es) const {
    const auto& mem = sec.chains[c];
    for (size_t k = 0; k < mem.size(); ++k) {
      es[k] = gemm_args(sec.ops[mem[k]]);
      es[k].kind = 0;
      es[k].bm = bm;
      es[k].bn = bn;
      es[k].bk = 32;
      es[k].ksplit = 1;
      es[k].nsplit = 1;
      es[k].kchunk = es[k].K;
      es[k].trans = sec.chain_flip[c][k] ^ trans;
    }
  }

  void run_chain(const Section& sec, int c) {
    const ccqk::GemmArgs& cp = sec.chain_plan[c];
    ccqk::GemmArgs es[ccqk::kMaxChain];
    chain_entries(sec, c, cp.trans, cp.bm, cp.bn, es);
    es[0].nsplit = cp.nsplit;
    es[0].kchunk = cp.kchunk;
    es[0].group = cp.group;
    HIPCHK(ccqk::launch_gemm_chain(stream, es, (int)sec.chains[c].size()));
  }

  // Grouped GEMM launches.  The ops of a compute section are placed on
  // levels: an op's level is one more than the highest level of the earlier
  // ops it depends on through memory (it reads what they write, or writes
  // what they read or write; tensors compared by address range).  Ops keep
  // their relative order inside a level, so every result is the one of the
  // sequential order.  On each level, tile GEMMs whose own tiles leave CUs
  // idle and that share a kernel instantiation run as one launch without
  // split-K.  CCQ_HIP_GROUP=0 keeps one launch per op in program order.
  bool group = true;
  // CCSDt triples residuals on the P space only (CCQ_HIP_PSPACE=0: full blocks)
  bool pspace = true;
  // measuring autotuner over the planner's best candidate plans (CCQ_HIP_TUNE=1);
  // with it, the sweep's products are also timed in FP32 storage, whose
  // float shadows it then allocates (CCQ_HIP_TUNE32=0: FP32 sweeps use the
  // FP64 plans)
  bool tune = false;
  bool tune32 = true;
  // point-group blocking: 0-based irrep of every correlated orbital (abelian
  // groups up to D2h, products by XOR); every tensor of the programs is
  // totally symmetric, so a product C = A B splits into the blocks of rows,
  // summed indices and columns of one irrep each (CCQ_HIP_SYM=0: dense)
  std::vector<int> orb_irrep;
  bool use_sym = false;
  int part_count = 0;
  bool dump_schedule = false;   // CCQ_HIP_SCHEDULE=1: levels and launches to stderr
  static constexpr int64_t kGroupTiles = 512;   // below this a GEMM joins groups
  // products with K >= group_min_k join groups; small register-direct
  // products (K <= 2048, M N <= 65536) join as 32 x 32 tile members
  // (CCQ_HIP_GROUP_MINK, CCQ_HIP_GROUP_REG=0; both cut launches of the
  // latency-bound sweeps: CCSDT 40 orbitals 7.70 -> 7.37 ms)
  int64_t group_min_k = 16;
  static constexpr int64_t kPartGroupMaxK = 1024;   // longest K of a part in a grouped launch
  bool group_reg = true;

  std::pair<uintptr_t, uintptr_t> span(int t) const {
    const Tensor& T = tensors[t];
    int64_t n = 1;
    for (int i = 0; i < T.rank; ++i) n *= T.dims[i];
    const uintptr_t b = reinterpret_cast<uintptr_t>(T.ptr);
    return {b, b + sizeof(double) * (uintptr_t)std::max<int64_t>(n, 1)};
  }

  // the op's plan as a member of a grouped launch: its tile, full K in
  // 32-deep slabs, no split
  static ccqk::GemmArgs group_plan(const Op& op) {
    if (op.has_gplan) return op.gplan;
    ccqk::GemmArgs g = op.plan;
    if (g.kind != 0) {
      // a register-direct product as a 32 x 32 tile member
      g.kind = 0;
      g.trans = 0;
      g.bm = g.bn = 32;
      g.ksplit = 1;
    }
    g.nsplit = 1;
    g.kchunk = g.K;
    if (g.kind == 0 && g.ksplit == 1) g.bk = 32;
    return g;
  }

  void group_gemms(bool dump) {
    for (auto& sec : sections) {
      if (sec.kind != Section::Compute) continue;
      // launch units: ops, with a chain as its head
      std::vector<int> idx;
      for (int i = 0; i < (int)sec.ops.size(); ++i)
        if (!sec.ops[i].fused && (sec.ops[i].chain < 0 || sec.ops[i].chain_head)) idx.push_back(i);
      if (!group) {
        for (int i : idx) {
          sec.launches.push_back({i});
          sec.launch_cls.push_back(sec.ops[i].chain >= 0 ? chain_cls(sec, sec.ops[i].chain) : sec.ops[i].cls);
          sec.launch_level.push_back((int)sec.launch_level.size());
        }
        continue;
      }
      const int n = (int)idx.size();
      std::vector<std::vector<std::pair<uintptr_t, uintptr_t>>> rd(n), wr(n);
      for (int q = 0; q < n; ++q) {
        const Op& h = sec.ops[idx[q]];
        std::vector<int> mem{idx[q]};
        if (h.chain >= 0) mem = sec.chains[h.chain];
        for (int i : mem) {
          const Op& op = sec.ops[i];
          rd[q].push_back(span(op.a));
          if (op.gemm) rd[q].push_back(span(op.b));
          if (op.fuse_a >= 0) rd[q].push_back(span(op.fuse_a));
        }
        if (!h.assign) rd[q].push_back(span(h.out));
        wr[q].push_back(span(h.out));
      }
      auto meet = [](const std::vector<std::pair<uintptr_t, uintptr_t>>& x,
                     const std::vector<std::pair<uintptr_t, uintptr_t>>& y) {
        for (auto& u : x)
          for (auto& v : y)
            if (u.first < v.second && v.first < u.second) return true;
        return false;
      };
      std::vector<int> level(n, 0);
      int top = 0;
      for (int q = 0; q < n; ++q) {
        const int pq = sec.ops[idx[q]].part_of;
        for (int r = 0; r < q; ++r) {
          // parts of one product write disjoint elements and read the same operands
          if (pq >= 0 && sec.ops[idx[r]].part_of == pq) continue;
          if (level[r] >= level[q] &&
              (meet(wr[r], rd[q]) || meet(wr[r], wr[q]) || meet(rd[r], wr[q])))
            level[q] = level[r] + 1;
        }
        top = std::max(top, level[q]);
      }
      for (int lv = 0; lv <= top; ++lv) {
        std::vector<int> here;
        for (int q = 0; q < n; ++q)
          if (level[q] == lv) here.push_back(idx[q]);
        std::map<int, std::vector<int>> group_at;   // first member -> members
        std::set<int> grouped;
        // the parts of one product (irrep blocks, P-space rectangles) share a
        // tile plan and run as grouped launches of up to kMaxGroup parts
        {
          std::map<int, std::vector<int>> sib;
          for (int i : here) {
            const Op& op = sec.ops[i];
            if (op.gemm && op.chain < 0 && op.part_of >= 0 && op.has_gplan) sib[op.part_of].push_back(i);
          }
          for (int i : here) {
            const Op& op = sec.ops[i];
            if (!op.gemm && op.part_of >= 0) sib[-2 - op.part_of].push_back(i);   // add parts
          }
          for (auto& kv : sib) {
            auto& v = kv.second;
            if (v.size() < 2) continue;
            const size_t cap = kv.first >= 0 ? (size_t)ccqk::kMaxGroup : (size_t)ccqk::kMaxPermGroup;
            for (size_t lo = 0; lo < v.size(); lo += cap) {
              const size_t hi = std::min(v.size(), lo + cap);
              std::vector<int> members(v.begin() + lo, v.begin() + hi);
              if (members.size() < 2) {
                continue;
              }
              group_at[*std::min_element(members.begin(), members.end())] = members;
              grouped.insert(members.begin(), members.end());
            }
          }
        }
        // candidates by kernel instantiation
        std::map<int64_t, std::vector<int>> cand;
        for (int i : here) {
          const Op& op = sec.ops[i];
          if (grouped.count(i)) continue;
          if (!op.gemm || op.chain >= 0 || op.K < group_min_k) continue;
          if (op.plan.kind != 0 && !(group_reg && op.K <= 2048 && op.M * op.N <= 65536)) continue;
          const ccqk::GemmArgs g = group_plan(op);
          if (ccqk::gemm_tiles(g) >= kGroupTiles) continue;
          const int64_t key = ccqk::gemm_group_key(g);
          if (key >= 0) cand[key].push_back(i);
        }
        // a launch lasts about as long as its longest workgroups: members of
        // one group have K within a factor 2 (runs of the K-sorted
        // candidates), longest K first, so that the workgroups dispatched
        // last are the short ones
        std::vector<std::vector<int>> runs;
        for (auto& kv : cand) {
          auto v = kv.second;
          std::stable_sort(v.begin(), v.end(),
                           [&](int x, int y) { return sec.ops[x].K > sec.ops[y].K; });
          for (size_t a = 0; a < v.size();) {
            size_t b = a + 1;
            while (b < v.size() && sec.ops[v[a]].K <= 2 * sec.ops[v[b]].K) ++b;
            runs.emplace_back(v.begin() + a, v.begin() + b);
            a = b;
          }
        }
        for (const auto& v : runs) {
          int64_t tiles = 0;
          for (int i : v) tiles += ccqk::gemm_tiles(group_plan(sec.ops[i]));
          if (v.size() < 2 || tiles < 256) continue;
          const int nch = (int)((v.size() + ccqk::kMaxGroup - 1) / ccqk::kMaxGroup);
          for (int c = 0; c < nch; ++c) {
            const size_t lo = v.size() * c / nch, hi = v.size() * (c + 1) / nch;
            if (hi - lo < 2) continue;
            std::vector<int> members(v.begin() + lo, v.begin() + hi);
            group_at[*std::min_element(members.begin(), members.end())] = members;
            grouped.insert(members.begin(), members.end());
          }
        }
        if (dump) {
          for (int i : here)
            std::fprintf(stderr, "[schedule] %s level %d op %d %s%s tiles %lld key %lld K %lld\n",
                         sec.name.c_str(), lv, i, sec.ops[i].cls.c_str(),
                         sec.ops[i].chain >= 0 ? " (chain head)" : "",
                         (long long)ccqk::gemm_tiles(group_plan(sec.ops[i])),
                         (long long)ccqk::gemm_group_key(group_plan(sec.ops[i])), (long long)unit_k(sec, i));
        }
        // the flat block copies of this level as one launch
        {
          std::vector<int> adds;
          for (int i : here) {
            const Op& op = sec.ops[i];
            if (!op.gemm && op.flat && !grouped.count(i)) adds.push_back(i);
          }
          for (size_t lo = 0; adds.size() >= 2 && lo + 1 < adds.size(); lo += ccqk::kMaxAddGroup) {
            const size_t hi = std::min(adds.size(), lo + (size_t)ccqk::kMaxAddGroup);
            if (hi - lo < 2) break;
            std::vector<int> members(adds.begin() + lo, adds.begin() + hi);
            group_at[members[0]] = members;
            grouped.insert(members.begin(), members.end());
          }
        }
        // the permuted adds of this level (their own tables each) as
        // perm_group launches; adds into the amplitudes stay alone
        {
          std::vector<int> adds;
          for (int i : here) {
            const Op& op = sec.ops[i];
            if (!op.gemm && !op.flat && op.part_of < 0 && !grouped.count(i) &&
                tensors[op.out].kind != TKind::Amp)
              adds.push_back(i);
          }
          for (size_t lo = 0; lo + 1 < adds.size(); lo += ccqk::kMaxPermGroup) {
            const size_t hi = std::min(adds.size(), lo + (size_t)ccqk::kMaxPermGroup);
            if (hi - lo < 2) break;
            std::vector<int> members(adds.begin() + lo, adds.begin() + hi);
            group_at[members[0]] = members;
            grouped.insert(members.begin(), members.end());
          }
        }
        // this level in op order, each group where its first member was
        for (int i : here) {
          auto it = group_at.find(i);
          if (it != group_at.end()) {
            sec.launches.push_back(it->second);
            sec.launch_cls.push_back(sec.ops[i].gemm ? group_cls(sec, it->second)
                                     : (sec.ops[i].part_of >= 0 || !sec.ops[i].flat ? "add:perm*" : "add:flat*") +
                                           std::to_string(it->second.size()));
            sec.launch_level.push_back(lv);
          } else if (!grouped.count(i)) {
            sec.launches.push_back({i});
            sec.launch_cls.push_back(sec.ops[i].chain >= 0 ? chain_cls(sec, sec.ops[i].chain) : sec.ops[i].cls);
            sec.launch_level.push_back(lv);
          }
        }
      }
    }
  }

  // K of a launch unit (a chain: its concatenated K)
  int64_t unit_k(const Section& sec, int i) const {
    const Op& op = sec.ops[i];
    if (op.chain < 0) return op.K;
    int64_t k = 0;
    for (int j : sec.chains[op.chain]) k += sec.ops[j].K;
    return k;
  }

  std::string group_cls(const Section& sec, const std::vector<int>& members) const {
    std::map<std::string, int> shapes;
    for (int i : members) {
      const Op& op = sec.ops[i];
      shapes[std::to_string(op.M) + "x" + std::to_string(op.N) + "x" + std::to_string(unit_k(sec, i))]++;
    }
    std::string t = "gemm:";
    bool first = true;
    for (auto& kv : shapes) {
      t += (first ? "" : "+") + std::to_string(kv.second) + "*" + kv.first;
      first = false;
    }
    const ccqk::GemmArgs g = group_plan(sec.ops[members[0]]);
    bool packed = false;
    for (int i : members) {
      const Op& op = sec.ops[i];
      packed = packed || op.pack_k || op.t_rowC2 >= 0 || op.t_colC2 >= 0;
    }
    return t + "/" + std::to_string(g.bm) + "x" + std::to_string(g.bn) + "k" + std::to_string(g.bk) +
           (g.trans ? "t" : "") + (packed ? "p" : "") + "g";
  }

  // -------------------------------------------------------------------------
  // execution

  ccqk::GemmArgs gemm_args(const Op& op) const {
    ccqk::GemmArgs g = run_prec == 2 && op.has_plan32 ? op.plan32 : op.plan;
    g.A = sp(op.a);
    g.rowA = tables32 + op.t_rowA;
    g.kA = tables32 + op.t_kA;
    g.B = sp(op.b);
    g.kB = tables32 + op.t_kB;
    g.colB = tables32 + op.t_colB;
    g.C = sp(op.out);
    g.rowC = tables32 + op.t_rowC;
    g.colC = tables32 + op.t_colC;
    g.alpha = op.coef;
    g.accumulate = op.assign ? 0 : 1;
    g.ws = cur_ws;
    g.prec = run_prec;
    if (op.fuse_a >= 0) {
      g.D = sp(op.fuse_a);
      g.rowD = tables32 + op.t_rowD;
      g.colD = tables32 + op.t_colD;
      g.dcoef = op.dcoef;
    }
    if (op.t_rowC2 >= 0) g.rowC2 = tables32 + op.t_rowC2;
    if (op.t_colC2 >= 0) g.colC2 = tables32 + op.t_colC2;
    return g;
  }

  void run_group(const Section& sec, const std::vector<int>& members) {
    ccqk::GemmArgs gs[ccqk::kMaxGroup];
    const int n = (int)members.size();
    if (n > ccqk::kMaxGroup) throw Fail("grouped launch too large");
    int64_t total = 0;
    for (int k = 0; k < n; ++k) total += ccqk::gemm_tiles(group_plan(sec.ops[members[k]]));
    for (int k = 0; k < n; ++k) {
      const Op& op = sec.ops[members[k]];
      const ccqk::GemmArgs gp = group_plan(op);
      gs[k] = gemm_args(op);
      gs[k].kind = gp.kind;
      gs[k].trans = gp.trans;
      gs[k].bm = gp.bm;
      gs[k].bn = gp.bn;
      gs[k].ksplit = gp.ksplit;
      gs[k].nsplit = gp.nsplit;
      gs[k].kchunk = gp.kchunk;
      gs[k].bk = gp.bk;
      // tile-row group of the XCD-aware order: about square blocks of the
      // member's tiles that one XCD receives (an eighth of the launch)
      const int64_t tiles = ccqk::gemm_tiles(gp);
      const int64_t tm = ((gp.trans ? op.N : op.M) + gp.bm - 1) / gp.bm;
      const double per_xcd = std::max(1.0, std::min((double)tiles, (double)total / 8.0));
      gs[k].group = (int)std::max<int64_t>(1, std::min<int64_t>(tm, std::lround(std::sqrt(per_xcd))));
    }
    HIPCHK(ccqk::launch_gemm_group(stream, gs, n));
  }

  void run_perm_group(const Section& sec, const std::vector<int>& members) {
    if (amp_in_fp64(members.size(), [&](size_t k) { return sec.ops[members[k]].out; },
                    [&] { run_perm_group(sec, members); }))
      return;
    ccqk::PermGroup pg{};
    if (members.size() > (size_t)ccqk::kMaxPermGroup) throw Fail("grouped permuted add too large");
    pg.n = (int)members.size();
    int64_t total = 0;
    for (int k = 0; k < pg.n; ++k) {
      const Op& op = sec.ops[members[k]];
      pg.A[k] = sp(op.a);
      pg.rowA[k] = tables32 + op.t_rowA;
      pg.colA[k] = tables32 + op.t_kA;
      pg.C[k] = sp(op.out);
      pg.rowC[k] = tables32 + op.t_rowC;
      pg.colC[k] = tables32 + op.t_colC;
      pg.N[k] = op.N;
      pg.alpha[k] = op.coef;
      pg.accumulate[k] = op.assign ? 0 : 1;
      pg.start[k] = total;
      total += op.M * op.N;
    }
    pg.start[pg.n] = total;
    HIPCHK(ccqk::launch_perm_group(stream, pg, run_prec == 2));
  }

  // float shadow of amplitude block b = its FP64 values
  void cast_block(const std::string& b) {
    HIPCHK(ccqk::launch_cast_f32(stream, tvec32 + blk_off.at(b), tvec.p + blk_off.at(b), blk_size.at(b)));
  }

  // In FP32-storage sweeps an add into the amplitudes (the closed-shell
  // permuted copies, which read amplitudes only) runs on the FP64 storage
  // and its output blocks are cast to their float shadows; returns whether
  // it ran the launch.
  template <typename OutOf, typename Run>
  bool amp_in_fp64(size_t n, OutOf out_of, Run run) {
    if (run_prec != 2) return false;
    std::set<std::string> blocks;
    for (size_t k = 0; k < n; ++k)
      if (tensors[out_of(k)].kind == TKind::Amp) blocks.insert(tensors[out_of(k)].base);
    if (blocks.empty()) return false;
    for (size_t k = 0; k < n; ++k)
      if (tensors[out_of(k)].kind != TKind::Amp) throw Fail("launch mixes adds into amplitudes and others");
    run_prec = 0;
    try {
      run();
    } catch (...) {
      run_prec = 2;
      throw;
    }
    run_prec = 2;
    for (const auto& b : blocks) cast_block(b);
    return true;
  }

  void run_add_group(const Section& sec, const std::vector<int>& members) {
    if (amp_in_fp64(members.size(), [&](size_t k) { return sec.ops[members[k]].out; },
                    [&] { run_add_group(sec, members); }))
      return;
    ccqk::AddGroup ag{};
    if (members.size() > (size_t)ccqk::kMaxAddGroup) throw Fail("grouped add launch too large");
    ag.n = (int)members.size();
    int64_t total = 0;
    for (int k = 0; k < ag.n; ++k) {
      const Op& op = sec.ops[members[k]];
      ag.A[k] = sp(op.a);
      ag.C[k] = sp(op.out);
      ag.alpha[k] = op.coef;
      ag.accumulate[k] = op.assign ? 0 : 1;
      ag.start[k] = total;
      total += op.M * op.N;
    }
    ag.start[ag.n] = total;
    HIPCHK(ccqk::launch_axpy_group(stream, ag, run_prec == 2));
  }

  void run_op(const Op& op) {
    if (op.fused) return;
    const int acc = op.assign ? 0 : 1;
    if (op.gemm) {
      HIPCHK(ccqk::launch_gemm(stream, gemm_args(op)));
    } else {
      if (amp_in_fp64(1, [&](size_t) { return op.out; }, [&] { run_op(op); })) return;
      const int io32 = run_prec == 2;
      const double* A = sp(op.a);
      double* C = sp(op.out);
      // the float form has no alignment condition
      const bool aligned =
          io32 || ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(C)) & 15) == 0;
      if (op.flat && aligned)
        HIPCHK(ccqk::launch_axpy_flat(stream, op.M * op.N, A, C, op.coef, acc, io32));
      else
        HIPCHK(ccqk::launch_perm_add(stream, op.M, op.N, A, tables32 + op.t_rowA, tables32 + op.t_kA, C,
                                     tables32 + op.t_rowC, tables32 + op.t_colC, op.coef, acc, io32));
    }
  }

  ccqk::UpdateParams update_params(const UpdateStep& u) const {
    const Tensor& R = tensors[u.resid];
    ccqk::UpdateParams P{};
    P.rank = R.rank;
    P.nimg = u.nimg;
    P.total = 1;
    const std::string& b = u.block;
    // positions: first half occupied, second half virtual; spin pattern by block name
    const int rank = R.rank, half = rank / 2;
    std::string spins = b.substr(2);  // e.g. "aab" from t3aab, "a" from t1a
    for (int d = 0; d < rank; ++d) {
      P.dims[d] = R.dims[d];
      P.strides[d] = R.strides[d];
      P.total *= R.dims[d];
      const int k = d % half;
      const bool beta = spins[k] == 'b';
      P.spin[d] = beta ? 1 : 0;
      P.isvir[d] = d >= half ? 1 : 0;
      P.occ[d] = beta ? ob : oa;
      P.group[d] = -1;
    }
    // antisymmetric groups: same-spin runs within the occupied / virtual halves
    for (int h = 0; h < 2; ++h)
      for (int k = 0; k + 1 < half; ++k)
        if (spins[k] == spins[k + 1]) {
          P.group[h * half + k] = h * 16 + (spins[k] == 'b');
          P.group[h * half + k + 1] = h * 16 + (spins[k] == 'b');
        }
    P.shift = shift;
    if (u.t_coc >= 0) {
      P.total = u.n_oc * u.n_vc;
      P.n_vc = std::max<int64_t>(u.n_vc, 1);
      if (itab) {
        P.coc = itab + u.t_coc;
        P.cvc = itab + u.t_cvc;
      }
    }
    if (u.masked) {
      P.mask = 1;
      P.vneed = u.vneed;
      P.vforbid = u.vforbid;
      P.occ_start = occ_start;
      P.vir_end[0] = vir_end[0];
      P.vir_end[1] = vir_end[1];
    }
    return P;
  }

  void run_update(const UpdateStep& u) {
    ccqk::UpdateParams P = update_params(u);
    if (!P.coc || !P.cvc) throw Fail("update step " + u.block + " has no canonical tuple tables");
    const Tensor& R = tensors[u.resid];
    if (u.left) {
      // left_jacobi_update (src/update_lcc.f90:208-280): L -= LH / (Hbar_aa - Hbar_ii - E_cor + shift)
      P.shift = shift - left_ecor;
      double* l = lvec.p + blk_off.at(dimkey(u.block));
      HIPCHK(ccqk::launch_update(stream, P, l, R.ptr, itab + u.dev_perm_off, dtab + u.dev_sign_off,
                                 tens("cct3:HD.a").ptr, tens("cct3:HD.b").ptr, lsq.p + u.sq_off));
      return;
    }
    double* t = tvec.p + blk_off.at(u.block);
    // FP32 storage: the residual is read as float; the updated block is then
    // cast to its shadow in one coalesced pass (cheaper than scattered float
    // stores beside the FP64 ones)
    P.u32 = run_prec == 2;
    HIPCHK(ccqk::launch_update(stream, P, t, sp(u.resid), itab + u.dev_perm_off, dtab + u.dev_sign_off,
                               fdiag_a.p, fdiag_b.p));
    if (run_prec == 2) cast_block(u.block);
  }

  const Tensor& tens(const std::string& id) const {
    auto it = tindex.find(id);
    if (it == tindex.end()) throw Fail("no tensor " + id);
    return tensors[it->second];
  }

  void run_copy(const std::string& dst, const std::string& src) {
    if (blk_size.at(dst) != blk_size.at(src)) throw Fail("copy between blocks of different size");
    HIPCHK(ccqk::launch_copy(stream, tvec.p + blk_off.at(dst), tvec.p + blk_off.at(src),
                             blk_size.at(src)));
    if (run_prec == 2) cast_block(dst);
  }

  // A sweep (and a left-CC step) is a fixed sequence of kernel launches on
  // fixed buffers; it is captured once into a hipGraph and replayed
  // (CCQ_HIP_GRAPH=0 disables).  key: the scalar baked into the launches.
  bool use_graph = true;
  struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    double key = 0.0;
  };
  std::map<std::string, Graph> graphs;

  static void drop_graph(Graph& G) {
    if (G.x) (void)hipGraphExecDestroy(G.x);
    if (G.g) (void)hipGraphDestroy(G.g);
    G.x = nullptr;
    G.g = nullptr;
  }

  // float shadows ready, and the float copy of t_vec refreshed unless it
  // was still valid at this call (cast on the stream, outside any graph)
  void prep_sweep32() {
    prep_f32();
    if (!t32_valid_at_call) HIPCHK(ccqk::launch_cast_f32(stream, tvec32, tvec.p, dev_t));
  }

  void run_graphed(const std::string& prog, double key = 0.0) {
    if (prog == "sweep" && prec == 2) prep_sweep32();
    if (!use_graph || profile) {
      run_program(prog);
      return;
    }
    // one graph per program and arithmetic (the launches differ)
    Graph& G = graphs[prog + ":p" + std::to_string(prec)];
    if (G.x && G.key != key) drop_graph(G);
    if (!G.x) {
      HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        run_program(prog);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIPCHK(hipStreamEndCapture(stream, &G.g));
      HIPCHK(hipGraphInstantiate(&G.x, G.g, nullptr, nullptr, 0));
      G.key = key;
    }
    HIPCHK(hipGraphLaunch(G.x, stream));
  }

  void run_sweep() { run_graphed("sweep"); }

  void sweep(int64_t nsec = -1) {
    if (prec == 2) prep_sweep32();
    run_program("sweep", nsec);
  }

  std::string cur_sec;   // section being run (profile attribution)

  // rough launch time (s) of launch k of a section: MFMA or HBM bound plus a
  // fixed cost
  double launch_cost(const Section& sec, size_t k) const {
    double fl = 0, by = 0;
    const auto& L = sec.launches[k];
    std::vector<int> mem = L;
    if (L.size() == 1 && sec.ops[L[0]].chain >= 0) mem = sec.chains[sec.ops[L[0]].chain];
    for (int i : mem) fl += sec.ops[i].flops, by += sec.ops[i].bytes;
    return std::max(fl / 40e12, by / 3e12) + 4e-6;
  }

  // The launches [a, b) of one dependency level (independent of each other:
  // no two write the same tensor or read what another writes) on up to
  // nstreams streams: longest first, each to the least loaded stream.  The
  // side streams wait for the main stream's work so far and the main stream
  // waits for theirs at the end, so the next level starts after all of this
  // one (also under graph capture, where this becomes a fork / join).
  template <typename F>
  void run_level(const Section& sec, size_t a, size_t b, F&& run_launch) {
    const int ns = std::min<int>(nstreams, (int)(b - a));
    std::vector<size_t> order;
    for (size_t k = a; k < b; ++k) order.push_back(k);
    std::vector<double> cost(b - a);
    for (size_t k = a; k < b; ++k) cost[k - a] = launch_cost(sec, k);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return cost[x - a] > cost[y - a]; });
    std::vector<double> load(ns, 0.0);
    std::vector<int> where(b - a, 0);
    for (size_t k : order) {
      const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      where[k - a] = s;
      load[s] += cost[k - a];
    }
    HIPCHK(hipEventRecord(fork_ev, stream));
    for (int s = 1; s < ns; ++s) HIPCHK(hipStreamWaitEvent(side[s - 1], fork_ev, 0));
    hipStream_t main = stream;
    for (size_t k = a; k < b; ++k) {
      const int s = where[k - a];
      stream = s ? side[s - 1] : main;
      cur_ws = s ? ws_side[s - 1].p : ws.p;
      try {
        run_launch(k);
      } catch (...) {
        stream = main;
        cur_ws = ws.p;
        throw;
      }
    }
    stream = main;
    cur_ws = ws.p;
    for (int s = 1; s < ns; ++s) {
      HIPCHK(hipEventRecord(join_ev[s - 1], side[s - 1]));
      HIPCHK(hipStreamWaitEvent(stream, join_ev[s - 1], 0));
    }
  }

  void run_program(const std::string& prog, int64_t nsec = -1) {
    if (!ints_ready) throw Fail("integrals not uploaded");
    run_prec = prog == "sweep" ? prec : std::min(prec, 1);
    if (run_prec == 2 && (!f32_ready || views32_stale)) throw Fail("FP32 shadows not prepared");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (profile) {
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
    }
    auto timed = [&](const std::string& cls, double fl, double by, auto&& fn) {
      if (!profile) {
        fn();
        return;
      }
      HIPCHK(hipEventRecord(e0, stream));
      fn();
      HIPCHK(hipEventRecord(e1, stream));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      auto& p = prof[cls];
      p.calls++;
      p.ms += ms;
      p.flops += fl;
      p.bytes += by;
      // the same launch, summed per section (reported as "section:<name>")
      auto& q = prof["section:" + cur_sec];
      q.calls++;
      q.ms += ms;
      q.flops += fl;
      q.bytes += by;
    };
    int64_t isec = 0;
    for (auto& sec : sections) {
      if (sec.prog != prog) continue;
      if (nsec >= 0 && isec++ >= nsec) break;
      cur_sec = sec.name;
      if (sec.kind == Section::Cct3) {
        timed("cct3:" + sec.upd.block, 0, 0, [&] { run_cct3(sec.upd); });
      } else if (sec.kind == Section::Compute) {
        if (sec.launches.empty()) {
          for (auto& op : sec.ops)
            if (!op.fused) timed(op.cls, op.flops, op.bytes, [&] { run_op(op); });
        } else {
          auto run_launch = [&](size_t k) {
            const auto& L = sec.launches[k];
            if (L.size() == 1 && sec.ops[L[0]].chain >= 0) {
              const int c = sec.ops[L[0]].chain;
              double fl = 0, by = 0;
              for (int i : sec.chains[c]) fl += sec.ops[i].flops, by += sec.ops[i].bytes;
              timed(sec.launch_cls[k], fl, by, [&] { run_chain(sec, c); });
              return;
            }
            if (L.size() == 1) {
              const Op& op = sec.ops[L[0]];
              timed(op.cls, op.flops, op.bytes, [&] { run_op(op); });
              return;
            }
            double fl = 0, by = 0;
            for (int i : L) fl += sec.ops[i].flops, by += sec.ops[i].bytes;
            if (sec.ops[L[0]].gemm)
              timed(sec.launch_cls[k], fl, by, [&] { run_group(sec, L); });
            else if (sec.ops[L[0]].part_of >= 0 || !sec.ops[L[0]].flat)
              timed(sec.launch_cls[k], fl, by, [&] { run_perm_group(sec, L); });
            else
              timed(sec.launch_cls[k], fl, by, [&] { run_add_group(sec, L); });
          };
          for (size_t a = 0; a < sec.launches.size();) {
            size_t b = a + 1;
            while (b < sec.launches.size() && sec.launch_level[b] == sec.launch_level[a]) ++b;
            if (b - a == 1 || nstreams == 1 || profile) {
              for (size_t k = a; k < b; ++k) run_launch(k);
            } else {
              run_level(sec, a, b, run_launch);
            }
            a = b;
          }
        }
      } else if (sec.kind == Section::Update) {
        const Tensor& R = tensors[sec.upd.resid];
        double n = 1;
        for (int d = 0; d < R.rank; ++d) n *= R.dims[d];
        timed("update:" + sec.upd.block, 0, 8.0 * 3 * n, [&] { run_update(sec.upd); });
      } else {
        run_copy(sec.dst, sec.src);
      }
    }
    if (profile) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
  }

  // diis_record of the right amplitudes also computes the next residuum's
  // partial sums (rpart); a residuum of the same iteration that is the very
  // next call of the C ABI uses them (api_calls counts the calls)
  int64_t api_calls = 0, rec_call = -10, rec_iter = -1;
  DevBuf rpart;

  double reduce_partials(int count, const double* src = nullptr) {
    host_partial.resize(count);
    HIPCHK(hipMemcpyAsync(host_partial.data(), src ? src : partial.p, sizeof(double) * count,
                          hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    double s = 0.0;
    for (int i = 0; i < count; ++i) s += host_partial[i];
    return s;
  }

  double energy() {
    // src/energy.f90:138-270, same eight pieces
    double* t1a = tvec.p + blk_off["t1a"];
    double* t1b = tvec.p + blk_off["t1b"];
    // each piece's block partials in its own slice, one copy to the host,
    // then the pieces summed in order
    HIPCHK(ccqk::launch_dot_strided(stream, oa, va, Fa.p + oa, nc, t1a, partial.p, kPartialBlocks));
    HIPCHK(ccqk::launch_dot_strided(stream, ob, vb, Fb.p + ob, nc, t1b, partial.p + kPartialBlocks,
                                    kPartialBlocks));
    struct Case {
      const double* V;
      int64_t o1, o2;
      const char* t2;
      double* x;
      double* y;
      double w2, w11;
    };
    Case cases[3] = {{Vaa.p, oa, oa, "t2aa", t1a, t1a, 0.25, 0.5},
                     {Vab.p, oa, ob, "t2ab", t1a, t1b, 1.0, 1.0},
                     {vbb_ptr, ob, ob, "t2bb", t1b, t1b, 0.25, 0.5}};
    int piece = 2;
    for (auto& c : cases) {
      ccqk::EnergyParams P{};
      P.V = c.V;
      P.vs0 = nc * nc * nc;
      P.vs1 = nc * nc;
      P.vs2 = nc;
      P.vs3 = 1;
      P.vo1 = c.o1;
      P.vo2 = c.o2;
      P.t2 = tvec.p + blk_off[c.t2];
      P.t1x = c.x;
      P.t1y = c.y;
      P.o1 = c.o1;
      P.o2 = c.o2;
      P.v1 = nc - c.o1;
      P.v2 = nc - c.o2;
      P.w2 = c.w2;
      P.w11 = c.w11;
      HIPCHK(ccqk::launch_energy_pairs(stream, P, partial.p + (int64_t)piece++ * kPartialBlocks,
                                       kPartialBlocks));
    }
    host_partial.resize((size_t)piece * kPartialBlocks);
    HIPCHK(hipMemcpyAsync(host_partial.data(), partial.p, sizeof(double) * piece * kPartialBlocks,
                          hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    double total = 0.0;
    for (int q = 0; q < piece; ++q) {
      double sp = 0.0;
      for (int b = 0; b < kPartialBlocks; ++b) sp += host_partial[(size_t)q * kPartialBlocks + b];
      total += sp;
    }
    return total;
  }

  // -------------------------------------------------------------------------
  // CC(t;3) tail

  bool hbar_ready = false;

  void add_inter(const std::string& id, std::vector<int64_t> d) {
    Tensor t;
    t.id = id;
    t.kind = TKind::Inter;
    t.rank = (int)d.size();
    for (int i = 0; i < t.rank; ++i) t.dims[i] = d[i];
    tindex[id] = (int)tensors.size();
    tensors.push_back(t);
  }

  // H-bar one-body diagonals [occ | vir] per spin and cal_d3's arrays
  // (src/mm_correct.f90:345-481), shapes [x][y][z] as there
  void add_tail_tensors() {
    add_inter("cct3:HD.a", {nc});
    add_inter("cct3:HD.b", {nc});
    add_inter("cct3:D3.a1", {va, oa, oa});
    add_inter("cct3:D3.a2", {va, va, oa});
    add_inter("cct3:D3.b1", {va, ob, oa});
    add_inter("cct3:D3.b2", {vb, va, oa});
    add_inter("cct3:D3.c1", {vb, ob, oa});
    add_inter("cct3:D3.c2", {vb, va, ob});
    add_inter("cct3:D3.d1", {vb, ob, ob});
    add_inter("cct3:D3.d2", {vb, vb, ob});
  }

  // hbar2 + add_twobody_hbar (src/hbar.f90:10-720) from the current T1, T2
  void hbar() {
    if (!tail) throw Fail("H-bar needs the CC(t;3) program (level 6)");
    run_program("hbar");
    auto diag = [&](const char* src, int64_t n, double* dst) {
      HIPCHK(ccqk::launch_diag(stream, tens(src).ptr, n, n, dst));
    };
    double* ha = tens("cct3:HD.a").ptr;
    double* hb = tens("cct3:HD.b").ptr;
    diag("cct3:HB1.oo.aa", oa, ha);
    diag("cct3:HB1.vv.aa", va, ha + oa);
    diag("cct3:HB1.oo.bb", ob, hb);
    diag("cct3:HB1.vv.bb", vb, hb + ob);
    HIPCHK(hipStreamSynchronize(stream));
    hbar_ready = true;
  }

  void left_init(double ecor) {
    if (!tail) throw Fail("left-CC needs the CC(t;3) program (level 6)");
    left_ecor = ecor;
    // ccq's l_vec starts from zero here
    HIPCHK(hipMemsetAsync(lvec.p, 0, sizeof(double) * l_size, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }

  void left_update() {
    if (!hbar_ready) throw Fail("H-bar not built (call ccq_hip_hbar first)");
    run_graphed("left", left_ecor);
  }

  // calculate_left_energy (src/energy.f90:103-136): |LH| + E_cor
  double left_energy() {
    host_partial.resize((size_t)lsq_n);
    HIPCHK(hipMemcpyAsync(host_partial.data(), lsq.p, sizeof(double) * lsq_n, hipMemcpyDeviceToHost,
                          stream));
    HIPCHK(hipStreamSynchronize(stream));
    double s = 0.0;
    for (int64_t i = 0; i < lsq_n; ++i) s += host_partial[i];
    return std::sqrt(s) + left_ecor;
  }

  // cal_d3 as gathered dot products: out[m] = scale sum_k X[xr[m] + xk[k]] Y[yr[m] + yk[k]]
  // over the bare integrals (C order [p][q][r][s], occupied first per spin) and T2
  void plan_d3() {
    std::vector<int64_t> host;
    const int64_t n1 = nc, n2 = nc * nc, n3 = nc * nc * nc;
    auto put = [&](const std::vector<int64_t>& v) {
      int64_t o = (int64_t)host.size();
      host.insert(host.end(), v.begin(), v.end());
      return o;
    };
    struct Spec {
      const char* out;
      const double* X;
      const char* t2;
      int64_t e0, e1, e2;     // output extents [x][y][z]
      int64_t K;
      std::function<int64_t(int64_t, int64_t, int64_t)> xr, yr;
      std::function<int64_t(int64_t)> xk, yk;
      double scale;
    };
    const double* Va = Vaa.p;
    const double* Vm = Vab.p;
    const double* Vb = vbb_ptr;
    // d3d1's e loop starts at orbital occ_a + 2 (src/mm_correct.f90:459)
    const int64_t e0 = std::min<int64_t>(vb, std::max<int64_t>(0, oa + 1 - ob));
    std::vector<Spec> specs = {
        // d3a1[a][j][i] = sum_e <ji||ea> t2aa[i][j][a][e]
        {"cct3:D3.a1", Va, "t2aa", va, oa, oa, va,
         [&](int64_t a, int64_t j, int64_t i) { return j * n3 + i * n2 + oa * n1 + oa + a; },
         [&](int64_t a, int64_t j, int64_t i) { return ((i * oa + j) * va + a) * va; },
         [&](int64_t e) { return e * n1; }, [&](int64_t e) { return e; }, 1.0},
        // d3a2[b][a][i] = -sum_m <mi||ba> t2aa[i][m][a][b]
        {"cct3:D3.a2", Va, "t2aa", va, va, oa, oa,
         [&](int64_t b, int64_t a, int64_t i) { return i * n2 + (oa + b) * n1 + oa + a; },
         [&](int64_t b, int64_t a, int64_t i) { return i * oa * va * va + a * va + b; },
         [&](int64_t m) { return m * n3; }, [&](int64_t m) { return m * va * va; }, -1.0},
        // d3b1[a][j][i] = sum_e <i j|a e>_ab t2ab[i][j][a][e]
        {"cct3:D3.b1", Vm, "t2ab", va, ob, oa, vb,
         [&](int64_t a, int64_t j, int64_t i) { return i * n3 + j * n2 + (oa + a) * n1 + ob; },
         [&](int64_t a, int64_t j, int64_t i) { return ((i * ob + j) * va + a) * vb; },
         [&](int64_t e) { return e; }, [&](int64_t e) { return e; }, 1.0},
        // d3b2[b][a][i] = -sum_m <i m|a b>_ab t2ab[i][m][a][b]
        {"cct3:D3.b2", Vm, "t2ab", vb, va, oa, ob,
         [&](int64_t b, int64_t a, int64_t i) { return i * n3 + (oa + a) * n1 + ob + b; },
         [&](int64_t b, int64_t a, int64_t i) { return i * ob * va * vb + a * vb + b; },
         [&](int64_t m) { return m * n2; }, [&](int64_t m) { return m * va * vb; }, -1.0},
        // d3c1[b][j][i] = sum_e <i j|e b>_ab t2ab[i][j][e][b]
        {"cct3:D3.c1", Vm, "t2ab", vb, ob, oa, va,
         [&](int64_t b, int64_t j, int64_t i) { return i * n3 + j * n2 + oa * n1 + ob + b; },
         [&](int64_t b, int64_t j, int64_t i) { return (i * ob + j) * va * vb + b; },
         [&](int64_t e) { return e * n1; }, [&](int64_t e) { return e * vb; }, 1.0},
        // d3c2[b][a][j] = -sum_m <m j|a b>_ab t2ab[m][j][a][b]
        {"cct3:D3.c2", Vm, "t2ab", vb, va, ob, oa,
         [&](int64_t b, int64_t a, int64_t j) { return j * n2 + (oa + a) * n1 + ob + b; },
         [&](int64_t b, int64_t a, int64_t j) { return j * va * vb + a * vb + b; },
         [&](int64_t m) { return m * n3; }, [&](int64_t m) { return m * ob * va * vb; }, -1.0},
        // d3d1[a][j][i] = sum_{e >= e0} <ji||ea> t2bb[i][j][a][e]
        {"cct3:D3.d1", Vb, "t2bb", vb, ob, ob, vb - e0,
         [&](int64_t a, int64_t j, int64_t i) { return j * n3 + i * n2 + (ob + e0) * n1 + ob + a; },
         [&](int64_t a, int64_t j, int64_t i) { return ((i * ob + j) * vb + a) * vb + e0; },
         [&](int64_t e) { return e * n1; }, [&](int64_t e) { return e; }, 1.0},
        // d3d2[b][a][i] = -sum_m <mi||ba> t2bb[i][m][a][b]
        {"cct3:D3.d2", Vb, "t2bb", vb, vb, ob, ob,
         [&](int64_t b, int64_t a, int64_t i) { return i * n2 + (ob + b) * n1 + ob + a; },
         [&](int64_t b, int64_t a, int64_t i) { return i * ob * vb * vb + a * vb + b; },
         [&](int64_t m) { return m * n3; }, [&](int64_t m) { return m * vb * vb; }, -1.0},
    };
    for (auto& sp : specs) {
      Had h{};
      h.out = tindex.at(sp.out);
      h.X = sp.X;
      h.Y = tvec.p + blk_off.at(sp.t2);
      h.M = sp.e0 * sp.e1 * sp.e2;
      h.K = sp.K;
      h.scale = sp.scale;
      std::vector<int64_t> xr, yr, xk, yk;
      for (int64_t x = 0; x < sp.e0; ++x)
        for (int64_t y = 0; y < sp.e1; ++y)
          for (int64_t z = 0; z < sp.e2; ++z) {
            xr.push_back(sp.xr(x, y, z));
            yr.push_back(sp.yr(x, y, z));
          }
      for (int64_t k = 0; k < sp.K; ++k) {
        xk.push_back(sp.xk(k));
        yk.push_back(sp.yk(k));
      }
      // every gathered element must lie inside its array
      const int64_t xlim = nc * nc * nc * nc, ylim = blk_size.at(sp.t2);
      int64_t xmax = 0, ymax = 0;
      for (auto v : xr) xmax = std::max(xmax, v);
      for (auto v : yr) ymax = std::max(ymax, v);
      int64_t xkm = 0, ykm = 0;
      for (auto v : xk) xkm = std::max(xkm, v);
      for (auto v : yk) ykm = std::max(ykm, v);
      if (h.M > 0 && h.K > 0 && (xmax + xkm >= xlim || ymax + ykm >= ylim))
        throw Fail(std::string("cal_d3 table out of range for ") + sp.out);
      h.xr = put(xr);
      h.yr = put(yr);
      h.xk = put(xk);
      h.yk = put(yk);
      hads.push_back(h);
    }
    HIPCHK(hipMalloc(&d3tab, sizeof(int64_t) * std::max<size_t>(host.size(), 1)));
    if (!host.empty())
      HIPCHK(hipMemcpy(d3tab, host.data(), sizeof(int64_t) * host.size(), hipMemcpyHostToDevice));
  }

  void run_cct3(const UpdateStep& u) {
    const Tensor& L = tensors[u.resid];
    const Tensor& M = tensors[u.m3];
    ccqk::Cct3Params P{};
    const std::string sp = u.block.substr(2);
    int cs = 0;
    for (char c : sp) cs += c == 'b';
    P.cs = cs;
    P.nimg = u.nimg;
    P.total = 1;
    for (int d = 0; d < 6; ++d) {
      P.dims[d] = L.dims[d];
      P.strides[d] = L.strides[d];
      P.total *= L.dims[d];
      if (M.dims[d] != L.dims[d] || M.strides[d] != L.strides[d]) throw Fail("L3 / M3 layouts differ");
    }
    P.oa = oa;
    P.ob = ob;
    P.va = va;
    P.vb = vb;
    P.occ_start = occ_start;
    P.vir_end[0] = vir_end[0];
    P.vir_end[1] = vir_end[1];
    P.fd[0] = fdiag_a.p;
    P.fd[1] = fdiag_b.p;
    P.hd[0] = tens("cct3:HD.a").ptr;
    P.hd[1] = tens("cct3:HD.b").ptr;
    P.ovov[0][0] = tens("cct3:HB.ovov.aaaa").ptr;
    P.ovov[1][0] = tens("cct3:HB.ovov.baba").ptr;
    P.ovov[0][1] = tens("cct3:HB.ovov.abab").ptr;
    P.ovov[1][1] = tens("cct3:HB.ovov.bbbb").ptr;
    const char* sfx[3] = {"aaaa", "abab", "bbbb"};
    for (int x = 0; x < 3; ++x) {
      P.oooo[x] = tens(std::string("cct3:HB.oooo.") + sfx[x]).ptr;
      P.vvvv[x] = tens(std::string("cct3:HB.vvvv.") + sfx[x]).ptr;
    }
    const char* d3n[8] = {"a1", "a2", "b1", "b2", "c1", "c2", "d1", "d2"};
    for (int x = 0; x < 8; ++x) P.d3[x] = tens(std::string("cct3:D3.") + d3n[x]).ptr;
    HIPCHK(ccqk::launch_cct3(stream, P, L.ptr, M.ptr, itab + u.dev_perm_off, dtab + u.dev_sign_off,
                             cpart.p + (int64_t)cs * kPartialBlocks * 4, kPartialBlocks));
  }

  // crcc23 (src/mm_correct.f90:11-343): e23[A..D]
  void cct3(double* e23) {
    if (!hbar_ready) throw Fail("H-bar not built (call ccq_hip_hbar first)");
    if (hads.empty()) plan_d3();
    for (auto& h : hads)
      HIPCHK(ccqk::launch_hadamard_sum(stream, h.M, h.K, h.X, d3tab + h.xr, d3tab + h.xk, h.Y,
                                       d3tab + h.yr, d3tab + h.yk, tensors[h.out].ptr, h.scale));
    HIPCHK(hipMemsetAsync(cpart.p, 0, sizeof(double) * cpart.n, stream));
    run_program("cct3");
    host_partial.resize((size_t)cpart.n);
    HIPCHK(hipMemcpyAsync(host_partial.data(), cpart.p, sizeof(double) * cpart.n,
                          hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int x = 0; x < 4; ++x) e23[x] = 0.0;
    for (int cs = 0; cs < 4; ++cs)
      for (int b = 0; b < kPartialBlocks; ++b)
        for (int x = 0; x < 4; ++x) e23[x] += host_partial[((size_t)cs * kPartialBlocks + b) * 4 + x];
  }

  // -------------------------------------------------------------------------
  // starting guess for T2 (MP2 or the learned MLP correction of it)

  float* guess_params = nullptr;

  std::vector<ccqk::GuessBlock> guess_blocks() {
    auto mk = [&](const char* blk, const double* V, int64_t o1, int64_t o2, int64_t v1, int64_t v2,
                  const double* f1, const double* f2) {
      ccqk::GuessBlock B{};
      B.V = V;
      B.nc = nc;
      B.o1 = o1;
      B.o2 = o2;
      B.v1 = v1;
      B.v2 = v2;
      B.va0 = (f1 == fdiag_a.p) ? oa : ob;
      B.vb0 = (f2 == fdiag_a.p) ? oa : ob;
      B.f1 = f1;
      B.f2 = f2;
      B.t = tvec.p + blk_off.at(blk);
      // 1 / rms of the MP2 amplitudes of the block
      const int64_t total = o1 * o2 * v1 * v2;
      HIPCHK(ccqk::launch_guess_msq(stream, B, partial.p, kPartialBlocks));
      const double s2 = reduce_partials(kPartialBlocks);
      B.inv_sm = (total > 0 && s2 > 0.0) ? 1.0 / std::sqrt(s2 / (double)total) : 0.0;
      return B;
    };
    std::vector<ccqk::GuessBlock> out = {mk("t2aa", Vaa.p, oa, oa, va, va, fdiag_a.p, fdiag_a.p),
                                         mk("t2ab", Vab.p, oa, ob, va, vb, fdiag_a.p, fdiag_b.p),
                                         mk("t2bb", vbb_ptr, ob, ob, vb, vb, fdiag_b.p, fdiag_b.p)};
    out[0].anti = out[2].anti = 1;
    return out;
  }

  // mode 0: T = 0; 1: T2 = MP2 amplitudes; 2: T2 = MP2 (1 + MLP); T1, T3, T4 = 0
  void guess(int64_t mode, const float* params, int64_t nparams) {
    if (!ints_ready) throw Fail("integrals not uploaded");
    if (mode < 0 || mode > 2) throw Fail("guess mode must be 0 (zero), 1 (MP2) or 2 (learned)");
    if (mode == 2 && (!params || nparams != ccqk::kGuessParams))
      throw Fail("learned guess needs " + std::to_string(ccqk::kGuessParams) + " parameters");
    HIPCHK(ccqk::launch_fill(stream, tvec.p, dev_t, 0.0));
    if (mode == 0) {
      HIPCHK(hipStreamSynchronize(stream));
      return;
    }
    if (mode == 2) {
      if (!guess_params) HIPCHK(hipMalloc(&guess_params, sizeof(float) * ccqk::kGuessParams));
      HIPCHK(hipMemcpyAsync(guess_params, params, sizeof(float) * ccqk::kGuessParams,
                            hipMemcpyHostToDevice, stream));
    }
    for (auto& B : guess_blocks())
      HIPCHK(ccqk::launch_guess_apply(stream, B, mode == 2 ? guess_params : nullptr));
    HIPCHK(hipStreamSynchronize(stream));
  }

  int64_t guess_data_size() const { return oa * oa * va * va + oa * ob * va * vb + ob * ob * vb * vb; }

  // features / targets of every T2 tuple at the current amplitudes
  void guess_data(float* X, float* M, float* Y, int64_t n) {
    if (!ints_ready) throw Fail("integrals not uploaded");
    if (n != guess_data_size()) throw Fail("guess data size mismatch");
    float *dX = nullptr, *dM = nullptr, *dY = nullptr;
    auto release = [&] {
      if (dX) (void)hipFree(dX);
      if (dM) (void)hipFree(dM);
      if (dY) (void)hipFree(dY);
    };
    try {
      HIPCHK(hipMalloc(&dX, sizeof(float) * std::max<int64_t>(n, 1) * ccqk::kGuessFeatures));
      HIPCHK(hipMalloc(&dM, sizeof(float) * std::max<int64_t>(n, 1)));
      HIPCHK(hipMalloc(&dY, sizeof(float) * std::max<int64_t>(n, 1)));
      int64_t off = 0;
      for (auto& B : guess_blocks()) {
        HIPCHK(ccqk::launch_guess_data(stream, B, dX + off * ccqk::kGuessFeatures, dM + off, dY + off));
        off += B.o1 * B.o2 * B.v1 * B.v2;
      }
      HIPCHK(hipStreamSynchronize(stream));
      HIPCHK(hipMemcpy(X, dX, sizeof(float) * n * ccqk::kGuessFeatures, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(M, dM, sizeof(float) * n, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(Y, dY, sizeof(float) * n, hipMemcpyDeviceToHost));
    } catch (...) {
      release();
      throw;
    }
    release();
  }

  // the DIIS vector the ring works on: t_vec, or l_vec in left mode
  int64_t diis_len() const { return left_mode ? l_size : vec_size; }

  int64_t slot_of(int64_t iter) const {
    int64_t s = iter % (diis + 1);
    if (s == 0) s = diis + 1;
    return s - 1;  // zero-based slot
  }
};

// ---------------------------------------------------------------------------
// helpers

namespace {

// LU with partial pivoting (the dgesv algorithm) for the small DIIS system.
bool solve_dense(std::vector<double> A, std::vector<double>& b, int n) {
  for (int k = 0; k < n; ++k) {
    int piv = k;
    double best = std::fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > best) best = std::fabs(A[i * n + k]), piv = i;
    if (best == 0.0) return false;
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[piv * n + j]);
      std::swap(b[k], b[piv]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] / A[k * n + k];
      A[i * n + k] = f;
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * b[j];
    b[i] = s / A[i * n + i];
  }
  return true;
}

template <class F>
// keep: the call does not write t_vec (energy, DIIS record, residuum, reads,
// switches), so the float copy of t_vec stays valid across it
int guarded(ccq_hip_ctx* ctx, F&& fn, bool keep = false) {
  if (!ctx) return 2;
  ++ctx->api_calls;
  ctx->t32_valid_at_call = ctx->t32_valid;
  if (!keep) ctx->t32_valid = false;
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return 1;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int ccq_hip_create(const ccq_hip_dims* dims, const ccq_hip_opts* opts, ccq_hip_ctx** out) {
  if (!dims || !opts || !out) return 2;
  *out = nullptr;
  std::unique_ptr<ccq_hip_ctx> ctx(new ccq_hip_ctx);
  int rc = guarded(ctx.get(), [&] {
    ctx->froz = dims->froz;
    ctx->occ_a = dims->occ_a;
    ctx->occ_b = dims->occ_b;
    ctx->orbs = dims->orbs;
    ctx->oa = dims->occ_a - dims->froz;
    ctx->ob = dims->occ_b - dims->froz;
    ctx->nc = dims->orbs - dims->froz;
    ctx->va = dims->orbs - dims->occ_a;
    ctx->vb = dims->orbs - dims->occ_b;
    if (ctx->oa < 0 || ctx->ob < 0 || ctx->va < 0 || ctx->vb < 0 || ctx->nc <= 0)
      throw Fail("inconsistent orbital counts");
    ctx->level = (int)opts->level;
    if (ctx->level < 2 || ctx->level > 6)
      throw Fail("level must be 2 (CCSD), 3 (CCSDT), 4 (CCSDTQ), 5 (CCSDt) or 6 (CC(t;3))");
    if (ctx->level >= 5) {
      // CC(t;3) = CCSDt, then H-bar, left-CCSD and crcc23 (src/calc_driver.f90:60-72)
      ctx->tail = ctx->level == 6;
      // CCSDt: CCSDT equations with active-space triples, no closed-shell
      // shortcuts (update_clusters_t3_opt, src/update_cc.F90:325-2191)
      if (opts->act_occ == -1 && opts->act_unocc == -1)
        throw Fail("CONFIGURATION ERROR: Active space methods require an active space");
      ctx->active = true;
      ctx->level = 3;
      ctx->act_occ = opts->act_occ;
      ctx->act_unocc = opts->act_unocc;
      // src/parser.f90:496-499
      const int64_t m1 = std::max(ctx->occ_b - ctx->act_occ, ctx->froz);
      const int64_t m2 = std::min(ctx->occ_a + ctx->act_unocc, ctx->orbs);
      ctx->occ_start = m1 - ctx->froz;
      ctx->vir_end[0] = std::max<int64_t>(0, m2 - ctx->occ_a);
      ctx->vir_end[1] = std::max<int64_t>(0, m2 - ctx->occ_b);
    }
    ctx->rhf = opts->rhf != 0 && !ctx->active;
    if (ctx->rhf && ctx->oa != ctx->ob) throw Fail("rhf requires occ_a == occ_b");
    ctx->shift = opts->shift;
    ctx->diis = (int)opts->diis_space;
    if (ctx->diis < 1 || ctx->diis > ccqk::kMaxDiis) throw Fail("diis_space must be in 1..12");
    ctx->device = (int)opts->device;
    if (const char* g = std::getenv("CCQ_HIP_GRAPH")) ctx->use_graph = std::string(g) != "0";
    if (const char* g = std::getenv("CCQ_HIP_FUSE")) ctx->fuse = std::string(g) != "0";
    if (const char* g = std::getenv("CCQ_HIP_GROUP")) ctx->group = std::string(g) != "0";
    if (const char* g = std::getenv("CCQ_HIP_PACK")) ctx->pack_pairs = std::string(g) != "0";
    if (const char* g = std::getenv("CCQ_HIP_CHAIN")) ctx->chain = std::string(g) != "0";
    if (const char* g = std::getenv("CCQ_HIP_VIEWPERM")) ctx->view_perm = std::string(g) != "0";
    if (const char* g = std::getenv("CCQ_HIP_SCHEDULE")) ctx->dump_schedule = std::string(g) == "1";
    if (const char* g = std::getenv("CCQ_HIP_PSPACE")) ctx->pspace = std::string(g) != "0";
    if (const char* g = std::getenv("CCQ_HIP_TUNE")) ctx->tune = std::string(g) == "1";
    if (const char* g = std::getenv("CCQ_HIP_TUNE32")) ctx->tune32 = std::string(g) != "0";
    if (const char* g = std::getenv("CCQ_HIP_GROUP_MINK")) ctx->group_min_k = std::max(1, std::atoi(g));
    if (const char* g = std::getenv("CCQ_HIP_GROUP_REG")) ctx->group_reg = std::string(g) != "0";
    if (const char* g = std::getenv("CCQ_HIP_MERGE")) ctx->merge_sections = std::string(g) != "0";
    if (const char* g = std::getenv("CCQ_HIP_STREAMS"))
      ctx->nstreams = std::max(1, std::min(ccq_hip_ctx::kMaxStreams, std::atoi(g)));
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->layout();
    const int64_t nc4 = ctx->nc * ctx->nc * ctx->nc * ctx->nc;
    ctx->Vaa.alloc(nc4);
    ctx->Vab.alloc(nc4);
    ctx->Fa.alloc(ctx->nc * ctx->nc);
    ctx->Fb.alloc(ctx->nc * ctx->nc);
    ctx->fdiag_a.alloc(ctx->nc);
    ctx->fdiag_b.alloc(ctx->nc);
    ctx->tvec.alloc(ctx->dev_t);
    HIPCHK(hipMemsetAsync(ctx->tvec.p, 0, sizeof(double) * ctx->dev_t, ctx->stream));
    ctx->ring.alloc((int64_t)(ctx->diis + 1) * ctx->vec_size);
    if (ctx->active) ctx->staging.alloc(ctx->ext_size);
    if (ctx->tail) {
      ctx->l_size = ctx->blk_off["t3aaa"];
      ctx->lvec.alloc(ctx->l_size);
      HIPCHK(hipMemsetAsync(ctx->lvec.p, 0, sizeof(double) * ctx->l_size, ctx->stream));
      ctx->cpart.alloc(4 * kPartialBlocks * 4);
      if (ctx->l_size > ctx->vec_size) throw Fail("DIIS ring too small for l_vec");
    }
    ctx->partial.alloc(kPartialBlocks * 128);
    if (ctx->level == 2) ctx->load_ir(ctx->rhf ? kIR_CCSD_RHF : kIR_CCSD_UHF);
    else if (ctx->active) {
      ctx->load_ir(kIR_CCSDT_ACT_UHF);
      if (ctx->tail) {
        ctx->load_ir(kIR_CCT3_UHF, "cct3:");
        ctx->add_tail_tensors();
      }
    }
    else if (ctx->level == 3) ctx->load_ir(ctx->rhf ? kIR_CCSDT_RHF : kIR_CCSDT_UHF);
    else {
#ifdef CCQ_HAVE_CCSDTQ
      ctx->load_ir(ctx->rhf ? kIR_CCSDTQ_RHF : kIR_CCSDTQ_UHF);
#else
      throw Fail("CCSDTQ program was not generated into this build");
#endif
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
  if (rc != 0) {
    fprintf(stderr, "ccq_hip_create: %s\n", ctx->err.c_str());
    return rc;
  }
  *out = ctx.release();
  return 0;
}

int ccq_hip_set_ints(ccq_hip_ctx* ctx, const double* f_a, const double* f_b, const double* v_aa,
                     const double* v_ab, const double* v_bb) {
  return guarded(ctx, [&] {
    const int64_t nc = ctx->nc, n3 = ctx->orbs, n0 = ctx->froz;
    const int64_t nc4 = nc * nc * nc * nc;
    // Fortran v(p,q,r,s) column-major -> C order [p][q][r][s]: full axis reversal
    std::vector<double> buf((size_t)nc4);
    auto rev4 = [&](const double* src, DevBuf& dst) {
      for (int64_t p = 0; p < nc; ++p)
        for (int64_t q = 0; q < nc; ++q)
          for (int64_t r = 0; r < nc; ++r) {
            double* o = buf.data() + ((p * nc + q) * nc + r) * nc;
            const double* s = src + p + nc * q + nc * nc * r;
            for (int64_t x = 0; x < nc; ++x) o[x] = s[x * nc * nc * nc];
          }
      HIPCHK(hipMemcpy(dst.p, buf.data(), sizeof(double) * nc4, hipMemcpyHostToDevice));
    };
    rev4(v_aa, ctx->Vaa);
    rev4(v_ab, ctx->Vab);
    // ccq builds v_bb identically to v_aa (src/integrals.f90:286-287); keep one
    // copy when they agree.
    const bool same = v_bb == v_aa || std::memcmp(v_bb, v_aa, sizeof(double) * nc4) == 0;
    if (!ctx->ints_ready) {
      if (same) {
        ctx->vbb_ptr = ctx->Vaa.p;
      } else {
        ctx->Vbb.alloc(nc4);
        ctx->vbb_ptr = ctx->Vbb.p;
      }
    } else if (same != (ctx->vbb_ptr == ctx->Vaa.p)) {
      throw Fail("v_bb changed between uploads");
    }
    if (!same) rev4(v_bb, ctx->Vbb);
    std::vector<double> fa(nc * nc), fb(nc * nc), da(nc), db(nc);
    for (int64_t p = 0; p < nc; ++p)
      for (int64_t q = 0; q < nc; ++q) {
        fa[p * nc + q] = f_a[(n0 + p) + n3 * (n0 + q)];
        fb[p * nc + q] = f_b[(n0 + p) + n3 * (n0 + q)];
      }
    for (int64_t p = 0; p < nc; ++p) da[p] = fa[p * nc + p], db[p] = fb[p * nc + p];
    HIPCHK(hipMemcpy(ctx->Fa.p, fa.data(), sizeof(double) * nc * nc, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->Fb.p, fb.data(), sizeof(double) * nc * nc, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->fdiag_a.p, da.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->fdiag_b.p, db.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
    if (!ctx->ints_ready) {
      ctx->resolve_tensors();
      ctx->plan();
    }
    ctx->gather_views();
    ctx->ints_ready = true;
  });
}

int ccq_hip_set_irreps(ccq_hip_ctx* ctx, const int64_t* irreps, int64_t n) {
  return guarded(ctx, [&] {
    if (ctx->ints_ready) throw Fail("ccq_hip_set_irreps must come before ccq_hip_set_ints");
    if (!irreps || n != ctx->orbs) throw Fail("ccq_hip_set_irreps: expected one irrep per orbital");
    ctx->orb_irrep.assign((size_t)ctx->nc, 0);
    bool any = false;
    for (int64_t p = 0; p < ctx->nc; ++p) {
      const int64_t v = irreps[ctx->froz + p];
      if (v < 1 || v > 8) throw Fail("ccq_hip_set_irreps: irreps are 1..8");
      ctx->orb_irrep[(size_t)p] = (int)(v - 1);
      any = any || v != 1;
    }
    const char* g = std::getenv("CCQ_HIP_SYM");
    ctx->use_sym = any && !(g && std::string(g) == "0");
  });
}

int64_t ccq_hip_t_size(const ccq_hip_ctx* ctx) { return ctx ? ctx->ext_size : -1; }

int ccq_hip_set_t(ccq_hip_ctx* ctx, const double* t_vec, int64_t t_size) {
  return guarded(ctx, [&] {
    if (t_size != ctx->ext_size) throw Fail("t_size mismatch");
    if (ctx->active) {
      HIPCHK(hipMemcpyAsync(ctx->staging.p, t_vec, sizeof(double) * ctx->ext_size,
                            hipMemcpyHostToDevice, ctx->stream));
      ctx->unpack(ctx->staging.p);
    } else {
      HIPCHK(hipMemcpyAsync(ctx->tvec.p, t_vec, sizeof(double) * std::min(ctx->dev_t, ctx->t_size),
                            hipMemcpyHostToDevice, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int ccq_hip_get_t(ccq_hip_ctx* ctx, double* t_vec, int64_t t_size) {
  return guarded(ctx, [&] {
    if (t_size != ctx->ext_size) throw Fail("t_size mismatch");
    if (ctx->active) {
      ctx->pack(ctx->staging.p);
      HIPCHK(hipMemcpyAsync(t_vec, ctx->staging.p, sizeof(double) * ctx->ext_size,
                            hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      return;
    }
    HIPCHK(hipMemcpyAsync(t_vec, ctx->tvec.p, sizeof(double) * std::min(ctx->dev_t, ctx->t_size),
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // blocks not resident (T3 for CCSD) are zero, as in ccq
    for (int64_t i = ctx->dev_t; i < ctx->t_size; ++i) t_vec[i] = 0.0;
  }, true);
}

int64_t ccq_hip_t4_size(const ccq_hip_ctx* ctx) {
  return ctx ? std::max<int64_t>(0, ctx->dev_t - ctx->t_size) : -1;
}

int ccq_hip_set_t4(ccq_hip_ctx* ctx, const double* t4, int64_t n) {
  return guarded(ctx, [&] {
    if (n != ctx->dev_t - ctx->t_size || n <= 0) throw Fail("t4 size mismatch");
    HIPCHK(hipMemcpyAsync(ctx->tvec.p + ctx->t_size, t4, sizeof(double) * n, hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int ccq_hip_get_t4(ccq_hip_ctx* ctx, double* t4, int64_t n) {
  return guarded(ctx, [&] {
    if (n != ctx->dev_t - ctx->t_size || n <= 0) throw Fail("t4 size mismatch");
    HIPCHK(hipMemcpyAsync(t4, ctx->tvec.p + ctx->t_size, sizeof(double) * n, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }, true);
}

// ---- ccq's amplitude files (Fortran unformatted records, restart.cpp) ----

int64_t ccq_hip_vec_size(const ccq_hip_ctx* ctx) { return ctx ? ctx->vec_size : -1; }

int ccq_hip_write_vecs(ccq_hip_ctx* ctx, int64_t iter, int64_t diis_space, const char* dir) {
  return guarded(ctx, [&] {
    if (iter < 1 || diis_space < 1) throw Fail("write_vecs: iter and diis_space must be >= 1");
    std::vector<double> t((size_t)ctx->ext_size);
    if (ccq_hip_get_t(ctx, t.data(), ctx->ext_size) != 0) throw Fail(ctx->err);
    int64_t slot = iter % (diis_space + 1);
    if (slot == 0) slot = diis_space + 1;
    std::string d = dir ? dir : "";
    if (!d.empty() && d.back() != '/') d += '/';
    for (const std::string& name : {std::string("t_vec.bin"),
                                    "cc_iter_mat-" + std::to_string(slot) + ".bin"})
      if (!ccqio::write_record(d + name, t.data(), ctx->vec_size))
        throw Fail("write_vecs: cannot write " + d + name);
  }, true);
}

int ccq_hip_read_vec(ccq_hip_ctx* ctx, const char* path) {
  return guarded(ctx, [&] {
    std::vector<double> t((size_t)ctx->ext_size);
    if (ccq_hip_get_t(ctx, t.data(), ctx->ext_size) != 0) throw Fail(ctx->err);
    std::string why;
    if (!ccqio::read_record(path ? path : "", t.data(), ctx->vec_size, why))
      throw Fail("read_vec: " + why);
    if (ccq_hip_set_t(ctx, t.data(), ctx->ext_size) != 0) throw Fail(ctx->err);
  });
}

namespace {
const char* const kT4Blocks[5] = {"t4aaaa", "t4aaab", "t4aabb", "t4abbb", "t4bbbb"};
const char* const kT4Files[5] = {"_4a.bin", "_4b.bin", "_4c.bin", "_4d.bin", "_4e.bin"};
}  // namespace

int ccq_hip_write_t4(ccq_hip_ctx* ctx, const char* bin_file) {
  return guarded(ctx, [&] {
    const int64_t n = ccq_hip_t4_size(ctx);
    if (n <= 0) throw Fail("write_t4: no T4 in this context (CCSDTQ only)");
    std::vector<double> t4((size_t)n);
    if (ccq_hip_get_t4(ctx, t4.data(), n) != 0) throw Fail(ctx->err);
    int64_t off = 0;
    for (int b = 0; b < 5; ++b) {
      const int64_t sz = ctx->blk_size.at(kT4Blocks[b]);
      const std::string path = std::string(bin_file ? bin_file : "") + kT4Files[b];
      if (!ccqio::write_record(path, t4.data() + off, sz)) throw Fail("write_t4: cannot write " + path);
      off += sz;
    }
  }, true);
}

int ccq_hip_read_t4(ccq_hip_ctx* ctx, const char* bin_file) {
  return guarded(ctx, [&] {
    const int64_t n = ccq_hip_t4_size(ctx);
    if (n <= 0) throw Fail("read_t4: no T4 in this context (CCSDTQ only)");
    std::vector<double> t4((size_t)n);
    int64_t off = 0;
    std::string why;
    for (int b = 0; b < 5; ++b) {
      const int64_t sz = ctx->blk_size.at(kT4Blocks[b]);
      const std::string path = std::string(bin_file ? bin_file : "") + kT4Files[b];
      if (!ccqio::read_record(path, t4.data() + off, sz, why)) throw Fail("read_t4: " + why);
      off += sz;
    }
    if (ccq_hip_set_t4(ctx, t4.data(), n) != 0) throw Fail(ctx->err);
  });
}

int ccq_hip_update(ccq_hip_ctx* ctx) {
  return guarded(ctx, [&] {
    ctx->run_sweep();
    ctx->t32_valid = ctx->prec == 2;
  });
}

int ccq_hip_energy(ccq_hip_ctx* ctx, double* e) {
  return guarded(ctx, [&] { *e = ctx->energy(); }, true);
}

int ccq_hip_diis_record(ccq_hip_ctx* ctx, int64_t iter) {
  return guarded(ctx, [&] {
    const int64_t s = ctx->slot_of(iter);
    if (ctx->left_mode) {
      HIPCHK(hipMemcpyAsync(ctx->ring.p + s * ctx->l_size, ctx->lvec.p, sizeof(double) * ctx->l_size,
                            hipMemcpyDeviceToDevice, ctx->stream));
      return;
    }
    if (ctx->active) {
      // the DIIS vector is ccq's packed t_vec (conv%vec_size = t_size)
      ctx->pack(ctx->ring.p + s * ctx->vec_size);
      return;
    }
    const double* prev = iter == 1 ? nullptr : ctx->ring.p + ctx->slot_of(iter - 1) * ctx->vec_size;
    if (!ctx->rpart.p) ctx->rpart.alloc(kPartialBlocks);
    HIPCHK(ccqk::launch_copy_diff_sq(ctx->stream, ctx->ring.p + s * ctx->vec_size, ctx->tvec.p, prev,
                                     ctx->vec_size, ctx->rpart.p, kPartialBlocks));
    ctx->rec_call = ctx->api_calls;
    ctx->rec_iter = iter;
  }, true);
}

int ccq_hip_residuum(ccq_hip_ctx* ctx, int64_t iter, double* res) {
  return guarded(ctx, [&] {
    if (!ctx->left_mode && !ctx->active && ctx->rec_iter == iter && ctx->rec_call == ctx->api_calls - 1) {
      *res = std::sqrt(std::sqrt(ctx->reduce_partials(kPartialBlocks, ctx->rpart.p)));
      return;
    }
    const int64_t len = ctx->diis_len();
    const double* cur = ctx->ring.p + ctx->slot_of(iter) * len;
    const double* prev = iter == 1 ? nullptr : ctx->ring.p + ctx->slot_of(iter - 1) * len;
    HIPCHK(ccqk::launch_diff_sq(ctx->stream, prev, cur, len, ctx->partial.p, kPartialBlocks));
    *res = std::sqrt(std::sqrt(ctx->reduce_partials(kPartialBlocks)));
  }, true);
}

int ccq_hip_diis_extrapolate(ccq_hip_ctx* ctx) {
  return guarded(ctx, [&] {
    const int n = ctx->diis;
    const int np = n * (n + 1) / 2;
    const int64_t len = ctx->diis_len();
    HIPCHK(ccqk::launch_diis_gram(ctx->stream, ctx->ring.p, len, n, ctx->partial.p,
                                  kPartialBlocks));
    ctx->host_partial.resize((size_t)kPartialBlocks * np);
    HIPCHK(hipMemcpyAsync(ctx->host_partial.data(), ctx->partial.p,
                          sizeof(double) * kPartialBlocks * np, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<double> g(np, 0.0);
    for (int b = 0; b < kPartialBlocks; ++b)
      for (int q = 0; q < np; ++q) g[q] += ctx->host_partial[(size_t)b * np + q];
    const int m = n + 1;
    std::vector<double> B((size_t)m * m, 0.0), c(m, 0.0);
    int q = 0;
    for (int x = 0; x < n; ++x)
      for (int y = x; y < n; ++y) {
        B[x * m + y] = g[q];
        B[y * m + x] = g[q];
        ++q;
      }
    for (int i = 0; i < m; ++i) {
      B[i * m + n] = -1.0;
      B[n * m + i] = -1.0;
    }
    c[n] = -1.0;
    if (!solve_dense(B, c, m)) throw Fail("calc_diis: DIIS error in DGESV.");
    ccqk::Coeffs cf{};
    for (int k = 0; k < n; ++k) cf.c[k] = c[k];
    if (ctx->left_mode) {
      HIPCHK(ccqk::launch_lincomb(ctx->stream, ctx->lvec.p, ctx->ring.p, len, n, cf));
    } else if (ctx->active) {
      HIPCHK(ccqk::launch_lincomb(ctx->stream, ctx->staging.p, ctx->ring.p, ctx->vec_size, n, cf));
      ctx->unpack(ctx->staging.p);
    } else {
      HIPCHK(ccqk::launch_lincomb(ctx->stream, ctx->tvec.p, ctx->ring.p, ctx->vec_size, n, cf));
    }
  });
}

int ccq_hip_hbar(ccq_hip_ctx* ctx) {
  return guarded(ctx, [&] { ctx->hbar(); });
}

int ccq_hip_left_init(ccq_hip_ctx* ctx, double e_cor) {
  return guarded(ctx, [&] { ctx->left_init(e_cor); });
}

int ccq_hip_left_update(ccq_hip_ctx* ctx) {
  return guarded(ctx, [&] { ctx->left_update(); });
}

int ccq_hip_left_energy(ccq_hip_ctx* ctx, double* e) {
  return guarded(ctx, [&] { *e = ctx->left_energy(); });
}

int ccq_hip_left_mode(ccq_hip_ctx* ctx, int64_t on) {
  return guarded(ctx, [&] {
    if (on && !ctx->tail) throw Fail("left mode needs the CC(t;3) program (level 6)");
    ctx->left_mode = on != 0;
  });
}

int64_t ccq_hip_l_size(const ccq_hip_ctx* ctx) { return ctx ? ctx->l_size : -1; }

int ccq_hip_set_l(ccq_hip_ctx* ctx, const double* l_vec, int64_t l_size) {
  return guarded(ctx, [&] {
    if (!ctx->tail || l_size != ctx->l_size) throw Fail("l_size mismatch");
    HIPCHK(hipMemcpyAsync(ctx->lvec.p, l_vec, sizeof(double) * l_size, hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int ccq_hip_get_l(ccq_hip_ctx* ctx, double* l_vec, int64_t l_size) {
  return guarded(ctx, [&] {
    if (!ctx->tail || l_size != ctx->l_size) throw Fail("l_size mismatch");
    HIPCHK(hipMemcpyAsync(l_vec, ctx->lvec.p, sizeof(double) * l_size, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }, true);
}

int ccq_hip_cct3(ccq_hip_ctx* ctx, double* e23) {
  return guarded(ctx, [&] { ctx->cct3(e23); });
}

int ccq_hip_guess(ccq_hip_ctx* ctx, int64_t mode, const float* params, int64_t nparams) {
  return guarded(ctx, [&] { ctx->guess(mode, params, nparams); });
}

int64_t ccq_hip_guess_data_size(const ccq_hip_ctx* ctx) { return ctx ? ctx->guess_data_size() : -1; }

int ccq_hip_guess_data(ccq_hip_ctx* ctx, float* X, float* M, float* Y, int64_t n) {
  return guarded(ctx, [&] { ctx->guess_data(X, M, Y, n); });
}

int ccq_hip_set_precision(ccq_hip_ctx* ctx, int64_t mode) {
  return guarded(ctx, [&] {
    if (mode < 0 || mode > 2)
      throw Fail("precision mode must be 0 (FP64), 1 (FP32 MFMA) or 2 (FP32 MFMA, FP32 storage)");
    ctx->prec = (int)mode;
  }, true);
}

int ccq_hip_profile(ccq_hip_ctx* ctx, int64_t enable) {
  return guarded(ctx, [&] {
    ctx->profile = enable != 0;
    ctx->prof.clear();
  }, true);
}

int ccq_hip_profile_report(ccq_hip_ctx* ctx, char* buf, int64_t buflen) {
  return guarded(ctx, [&] {
    std::ostringstream os;
    for (auto& kv : ctx->prof)
      os << kv.first << " " << kv.second.calls << " " << kv.second.ms << " " << kv.second.flops
         << " " << kv.second.bytes << "\n";
    std::string s = os.str();
    if ((int64_t)s.size() + 1 > buflen) throw Fail("profile buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
  }, true);
}

int ccq_hip_mfma_ceiling(int64_t kind, int64_t blocks, int64_t iters, double* seconds, double* flops) {
  if (kind < 0 || kind > 2 || blocks <= 0 || blocks > (1 << 20) || iters <= 0 || !seconds || !flops)
    return (int)hipErrorInvalidValue;
  double* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&out, sizeof(double) * 256 * blocks);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  double fl = 0.0;
  if (e == hipSuccess) e = ccqk::launch_mfma_ceiling(nullptr, (int)kind, blocks, iters, out, &fl);
  if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
  if (e == hipSuccess) e = ccqk::launch_mfma_ceiling(nullptr, (int)kind, blocks, iters, out, &fl);
  if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) {
    *seconds = ms * 1e-3;
    *flops = fl;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (out) (void)hipFree(out);
  return (int)e;
}

int ccq_hip_sync(ccq_hip_ctx* ctx) {
  return guarded(ctx, [&] { HIPCHK(hipStreamSynchronize(ctx->stream)); }, true);
}

int ccq_hip_debug_sweep(ccq_hip_ctx* ctx, int64_t nsections) {
  return guarded(ctx, [&] { ctx->sweep(nsections); });
}

int ccq_hip_debug_read(ccq_hip_ctx* ctx, const char* tensor_id, double* out, int64_t n) {
  return guarded(ctx, [&] {
    auto it = ctx->tindex.find(tensor_id);
    if (it == ctx->tindex.end()) throw Fail(std::string("no tensor ") + tensor_id);
    const Tensor& t = ctx->tensors[it->second];
    if (t.kind == TKind::View) throw Fail("views are not readable");
    int64_t total = 1;
    for (int d = 0; d < t.rank; ++d) total *= t.dims[d];
    if (n != total) throw Fail("debug_read size mismatch");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(out, t.ptr, sizeof(double) * total, hipMemcpyDeviceToHost));
  }, true);
}

const char* ccq_hip_last_error(const ccq_hip_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ccq_hip_destroy(ccq_hip_ctx* ctx) {
  if (!ctx) return 0;
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return 0;
}

}  // extern "C"
