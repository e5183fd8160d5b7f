// Host model of the FP32 LU-SGS sweep's synchronisation (test
// infrastructure: compiled and run by tests/test_lu32_protocol_model.py on
// the CPU, also under ThreadSanitizer; nothing of the library links it).
//
// What is modelled.  One direction of the blocked sweep of
// turns2d_urans_git_amd/csrc/ot2d_kernels.hip (k_lu_sweep32), scalar:
//   s(j,k) = u(j,k) + a(j,k) s(j-1,k) + b(j,k) s(j,k-1)
// (src/lhs.f:616-714 with the 4x4 operators replaced by scalars; the
// arithmetic does not matter here, the order of the hand-offs does).  The
// rows are cut into W blocks of R rows; a workgroup is one consumer thread
// (the computing wave: lane l of block w owns row R*w+l and reaches column
// t-l at step t) and NP producer threads (producer p fills the LDS ring
// slots of steps p, p+NP, ...).  Inside a workgroup:
//   ring slot t % S  records of step t, the fill value, tag = t+1
//                    (release store after the records)
//   consumed         = t+1 once the consumer has read slot t % S; a
//                    producer refills the slot of step t only when
//                    consumed >= t-S+1
// The LDS executes one wave's operations in issue order; the model states
// that with release/acquire on the tag and on `consumed`, which is what
// ThreadSanitizer checks the slot reuse against.  Across workgroups, block
// w needs block w-1's last-lane result of column t, which that block
// produced at its step t+R-1 (the mail).
//
// Two protocols:
//   sentinel (the earlier kernel): the mail is the last-lane entry of the
//     result array H; a separate prep kernel resets it to an EMPTY pattern
//     with plain stores before every sweep, and a reader accepts any
//     non-EMPTY value.  Workgroup x is block x (dispatch order).
//   tagged (the current kernel): the mail is 8-byte words {value, sweep
//     tag} (the kernel keeps them in its result rows, two words per 16-byte
//     store), written only by the sweep; a reader accepts a word only with
//     the running sweep's tag.  A workgroup takes its block from
//     an atomic ticket counter (ticket n: block n % W, sweep n / W, tag
//     n / W + 1), so it only ever waits on a workgroup that has started.
//
// Memory model.  Device memory is an array of std::atomic words.  A plain
// store made by a kernel goes into the write-back cache of the XCD its
// workgroup runs on and reaches memory when that cache is written back.
// With --late-writeback the write-back of a finished kernel's dirty lines
// may still be in flight when the next kernel starts (a delayed write-back
// across the kernel boundary: the one assumption the sentinel protocol
// needs and the tagged protocol does not).  System-scope loads and stores
// (the mail, in both protocols) go straight to memory.  Workgroups are
// dispatched round robin over XCDs, each XCD in order, into a limited
// number of CU slots shared by all streams.
//
// Checks, per sweep: every result of every instance equals a sequential
// evaluation in the same operation order, bit for bit, with no NaN; every
// wait ends within the wait limit (a wait that does not is what the kernel
// reports as status 3).  Exit status 0 when all hold, 1 otherwise; the
// counts are printed.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

struct Cfg {
  int ni = 24, nk = 11, R = 4, S = 4, NP = 2;
  int nsweep = 60, streams = 3, ninst = 2;
  int xcds = 2, slots = 3;            // CU slots per XCD, shared by all streams
  bool tagged = true;
  bool late_wb = false;
  bool plant_ring_bug = false;        // refill a slot one step early (TSan must flag it)
  int wait_ms = 400;                  // bounded waits (the kernel's 0.5 s)
  unsigned seed = 1;
};

const uint32_t EMPTY = 0x7fa5a5a5u;   // signalling-NaN pattern of the sentinel protocol

uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
float fval(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// random short pauses: the interleavings a loaded GPU produces
thread_local std::mt19937 rng_tl{12345};
void jitter(int pct) {
  if ((int)(rng_tl() % 100) < pct) {
    if (rng_tl() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng_tl() % 60));
    else std::this_thread::yield();
  }
}

// ------------------------------------------------------------ device memory
// Plain stores of a kernel are held per XCD until written back.
struct XcdCache {
  std::mutex mu;
  std::vector<std::pair<std::atomic<uint32_t>*, uint32_t>> dirty;
};

struct Memory {
  std::vector<XcdCache> xcd;
  std::mutex wb_mu;
  std::condition_variable wb_cv;
  std::deque<std::vector<std::pair<std::atomic<uint32_t>*, uint32_t>>> inflight;  // write-backs in flight
  bool stop = false;
  std::thread wb_thread;
  explicit Memory(int nx) : xcd(nx) {}
  void plain_store(int x, std::atomic<uint32_t>* p, uint32_t v) {
    std::lock_guard<std::mutex> g(xcd[x].mu);
    xcd[x].dirty.emplace_back(p, v);
  }
  // kernel end: every XCD's dirty lines are written back, at once or (late
  // write-back) by a background agent after a short delay
  void kernel_end(bool late) {
    std::vector<std::pair<std::atomic<uint32_t>*, uint32_t>> all;
    for (auto& c : xcd) {
      std::lock_guard<std::mutex> g(c.mu);
      all.insert(all.end(), c.dirty.begin(), c.dirty.end());
      c.dirty.clear();
    }
    if (all.empty()) return;
    if (!late) {
      for (auto& d : all) d.first->store(d.second, std::memory_order_release);
      return;
    }
    std::lock_guard<std::mutex> g(wb_mu);
    inflight.push_back(std::move(all));
    wb_cv.notify_one();
  }
  void start_writeback_agent() {
    wb_thread = std::thread([this] {
      std::mt19937 r(777);
      for (;;) {
        std::vector<std::pair<std::atomic<uint32_t>*, uint32_t>> b;
        {
          std::unique_lock<std::mutex> g(wb_mu);
          wb_cv.wait(g, [&] { return stop || !inflight.empty(); });
          if (inflight.empty() && stop) return;
          b = std::move(inflight.front());
          inflight.pop_front();
        }
        std::this_thread::sleep_for(std::chrono::microseconds(50 + r() % 400));
        for (auto& d : b) d.first->store(d.second, std::memory_order_release);
      }
    });
  }
  void stop_agent() {
    if (!wb_thread.joinable()) return;
    {
      std::lock_guard<std::mutex> g(wb_mu);
      stop = true;
    }
    wb_cv.notify_all();
    wb_thread.join();
  }
};

// ------------------------------------------------------------- one instance
struct Instance {
  int ni, nk, R, W, NT;
  std::vector<float> u, a, b;                 // operator of the current sweep (written by the prep)
  std::vector<std::atomic<uint32_t>> H;       // results, (block, step, lane)
  std::vector<std::atomic<uint64_t>> mail;    // tagged protocol: (block, step)
  std::atomic<uint64_t> ticket{0};
  Instance(int ni_, int nk_, int R_) : ni(ni_), nk(nk_), R(R_), W((nk_ + R_ - 1) / R_), NT(ni_ + R_ - 1),
      u((size_t)ni_ * nk_), a((size_t)ni_ * nk_), b((size_t)ni_ * nk_), H((size_t)W * NT * R_), mail((size_t)W * NT) {
    for (auto& h : H) h.store(0);
    for (auto& m : mail) m.store(0);
  }
  float& U(int j, int k) { return u[(size_t)k * ni + j]; }
  float& A(int j, int k) { return a[(size_t)k * ni + j]; }
  float& B(int j, int k) { return b[(size_t)k * ni + j]; }
  std::atomic<uint32_t>& Hat(int w, int t, int l) { return H[((size_t)w * NT + t) * R + l]; }
};

struct Counts {
  std::atomic<long> sweeps{0}, wrong{0}, nan{0}, timeouts{0}, stale_accepts{0};
};

// ------------------------------------------------------------------ one workgroup
struct Ring {
  std::vector<float> rec;          // [S][3R] u, a, b of each lane
  std::vector<float> fill;         // [S]
  std::vector<std::atomic<int>> tag;
  std::atomic<int> consumed{0};
  std::atomic<int> abort{0};
  explicit Ring(int S, int R) : rec((size_t)S * 3 * R), fill(S), tag(S) {
    for (auto& t : tag) t.store(0);
  }
};

bool spin_until(const std::function<bool()>& ok, Ring& rg, int wait_ms) {
  auto t0 = Clock::now();
  while (!ok()) {
    if (rg.abort.load(std::memory_order_relaxed)) return false;
    if (Clock::now() - t0 > std::chrono::milliseconds(wait_ms)) {
      rg.abort.store(1);
      return false;
    }
    std::this_thread::yield();
  }
  return true;
}

// one workgroup of one sweep of one instance; `slot_x` is the linear
// workgroup index within the instance (its blockIdx.x)
void run_workgroup(const Cfg& cfg, Instance& I, Memory& mem, int xcd, int slot_x, uint32_t sentinel_expect_tag,
                   const std::vector<float>& sref, Counts& cnt) {
  (void)sentinel_expect_tag;
  int w;
  uint32_t tag = 0;
  if (cfg.tagged) {
    const uint64_t n = I.ticket.fetch_add(1, std::memory_order_relaxed);
    w = (int)(n % (uint64_t)I.W);
    tag = (uint32_t)(n / (uint64_t)I.W) + 1u;
  } else {
    w = slot_x;
  }
  const int R = cfg.R, S = cfg.S, NP = cfg.NP, NT = I.NT;
  Ring rg(S, R);
  const bool has_in = w > 0;
  // the neighbour's mail for column c (its step c+R-1, last lane)
  auto read_mail = [&](int c, uint32_t& v, bool& stale_wrong) -> bool {
    if (cfg.tagged) {
      std::atomic<uint64_t>& m = I.mail[(size_t)(w - 1) * NT + c + R - 1];
      auto t0 = Clock::now();
      for (;;) {
        uint64_t x = m.load(std::memory_order_relaxed);   // system-scope atomic load
        if ((uint32_t)(x >> 32) == tag) {
          v = (uint32_t)x;
          return true;
        }
        if (rg.abort.load()) return false;
        if (Clock::now() - t0 > std::chrono::milliseconds(cfg.wait_ms)) {
          rg.abort.store(1);
          return false;
        }
        std::this_thread::yield();
      }
    } else {
      std::atomic<uint32_t>& m = I.Hat(w - 1, c + R - 1, R - 1);
      auto t0 = Clock::now();
      for (;;) {
        uint32_t x = m.load(std::memory_order_relaxed);   // system-scope load
        if (x != EMPTY) {
          v = x;
          // the value the sequential sweep has there: anything else is a
          // stale entry taken for the current one
          const float want = sref[(size_t)(R * (w - 1) + R - 1) * I.ni + c];
          stale_wrong = fbits(want) != x;
          return true;
        }
        if (rg.abort.load()) return false;
        if (Clock::now() - t0 > std::chrono::milliseconds(cfg.wait_ms)) {
          rg.abort.store(1);
          return false;
        }
        std::this_thread::yield();
      }
    }
  };
  std::vector<std::thread> prod;
  for (int p = 0; p < NP; ++p) {
    prod.emplace_back([&, p] {
      rng_tl.seed(cfg.seed * 1000003u + (unsigned)(w * 131 + p * 7 + xcd));
      for (int t = p; t < NT; t += NP) {
        jitter(20);
        float rec[3 * 64];
        for (int l = 0; l < R; ++l) {
          const int j = t - l, k = R * w + l;
          const bool in = j >= 0 && j < I.ni && k < I.nk;
          rec[l] = in ? I.U(j, k) : 0.f;
          rec[R + l] = in ? I.A(j, k) : 0.f;
          rec[2 * R + l] = in ? I.B(j, k) : 0.f;
        }
        uint32_t fill = 0;
        if (has_in && t < I.ni) {
          bool stale = false;
          if (!read_mail(t, fill, stale)) return;
          if (stale) cnt.stale_accepts++;
        }
        const int need = cfg.plant_ring_bug ? t - S : t - S + 1;
        if (!spin_until([&] { return rg.consumed.load(std::memory_order_acquire) >= need; }, rg, cfg.wait_ms))
          return;
        const int sl = t % S;
        for (int q = 0; q < 3 * R; ++q) rg.rec[(size_t)sl * 3 * R + q] = rec[q];
        rg.fill[sl] = fval(fill);
        rg.tag[sl].store(t + 1, std::memory_order_release);
      }
    });
  }
  // consumer (the computing wave)
  {
    rng_tl.seed(cfg.seed * 7919u + (unsigned)(w * 17 + xcd));
    std::vector<float> sA(R, 0.f), kout(R, 0.f), sn(R);
    const bool has_out = w < I.W - 1;
    bool ok = true;
    for (int t = 0; t < NT && ok; ++t) {
      jitter(10);
      const int sl = t % S;
      ok = spin_until([&] { return rg.tag[sl].load(std::memory_order_acquire) >= t + 1; }, rg, cfg.wait_ms);
      if (!ok) break;
      float rec[3 * 64], fill;
      for (int q = 0; q < 3 * R; ++q) rec[q] = rg.rec[(size_t)sl * 3 * R + q];
      fill = rg.fill[sl];
      rg.consumed.store(t + 1, std::memory_order_release);
      for (int l = 0; l < R; ++l) {
        const float kin = l == 0 ? fill : kout[l - 1];
        float p = std::fmaf(rec[R + l], sA[l], rec[l]);
        sn[l] = std::fmaf(rec[2 * R + l], kin, p);
      }
      for (int l = 0; l < R; ++l) {
        // result store (write-back cached in the sentinel protocol's time
        // only through memory: model it as a system-scope store, as the
        // kernel's result rows were)
        I.Hat(w, t, l).store(fbits(sn[l]), std::memory_order_relaxed);
        sA[l] = sn[l];
        kout[l] = sn[l];
      }
      if (cfg.tagged && has_out)
        I.mail[(size_t)w * NT + t].store(((uint64_t)tag << 32) | fbits(sn[R - 1]), std::memory_order_relaxed);
    }
    if (!ok) cnt.timeouts++;
  }
  for (auto& th : prod) th.join();
}

// -------------------------------------------------------------- dispatcher
struct Dispatcher {
  std::mutex mu;
  std::condition_variable cv;
  std::vector<int> free_slots;
  explicit Dispatcher(int nx, int slots) : free_slots(nx, slots) {}
  void acquire(int x) {
    std::unique_lock<std::mutex> g(mu);
    cv.wait(g, [&] { return free_slots[x] > 0; });
    --free_slots[x];
  }
  void release(int x) {
    {
      std::lock_guard<std::mutex> g(mu);
      ++free_slots[x];
    }
    cv.notify_all();
  }
};

void sequential(Instance& I, std::vector<float>& s) {
  s.assign((size_t)I.ni * I.nk, 0.f);
  for (int k = 0; k < I.nk; ++k)
    for (int j = 0; j < I.ni; ++j) {
      const float sj = j > 0 ? s[(size_t)k * I.ni + j - 1] : 0.f;
      const float sk = k > 0 ? s[(size_t)(k - 1) * I.ni + j] : 0.f;
      float p = std::fmaf(I.A(j, k), sj, I.U(j, k));
      s[(size_t)k * I.ni + j] = std::fmaf(I.B(j, k), sk, p);
    }
}

// one stream: nsweep launches (prep, then the sweep) over its instances
void run_stream(const Cfg& cfg, int sid, Memory& mem, Dispatcher& disp, std::atomic<int>& linear_base, Counts& cnt) {
  std::vector<std::unique_ptr<Instance>> inst;
  for (int i = 0; i < cfg.ninst; ++i) inst.emplace_back(new Instance(cfg.ni, cfg.nk - i, cfg.R));
  std::mt19937 r(cfg.seed * 31u + (unsigned)sid);
  std::uniform_real_distribution<float> U(-1.f, 1.f), A(0.05f, 0.45f);
  std::vector<std::vector<float>> sref(cfg.ninst);
  for (int sw = 0; sw < cfg.nsweep; ++sw) {
    // ---- prep kernel: this sweep's operator; the sentinel protocol also
    // resets the mail (plain stores, from the XCD the prep workgroup runs on)
    for (int i = 0; i < cfg.ninst; ++i) {
      Instance& I = *inst[i];
      for (auto& v : I.u) v = U(r);
      for (auto& v : I.a) v = A(r);
      for (auto& v : I.b) v = A(r);
      sequential(I, sref[i]);
      if (!cfg.tagged)
        for (int w = 0; w < I.W; ++w)
          for (int c = 0; c < I.ni; ++c) mem.plain_store((w + c) % cfg.xcds, &I.Hat(w, c + cfg.R - 1, cfg.R - 1), EMPTY);
    }
    mem.kernel_end(cfg.late_wb);
    // ---- sweep kernel: grid (W, ninst); linear workgroup n goes to XCD
    // (base + n) % xcds, each XCD dispatching in order
    const int wmax = inst[0]->W;
    const int base = linear_base.fetch_add(wmax * cfg.ninst);
    std::vector<std::vector<std::pair<int, int>>> q(cfg.xcds);   // (instance, blockIdx.x)
    for (int i = 0; i < cfg.ninst; ++i)
      for (int x = 0; x < wmax; ++x) q[(base + i * wmax + x) % cfg.xcds].emplace_back(i, x);
    std::vector<std::thread> xthreads;
    for (int x = 0; x < cfg.xcds; ++x) {
      xthreads.emplace_back([&, x] {
        std::vector<std::thread> wgs;
        for (auto& e : q[x]) {
          disp.acquire(x);
          wgs.emplace_back([&, x, e] {
            Instance& I = *inst[e.first];
            if (e.second < I.W) run_workgroup(cfg, I, mem, x, e.second, 0, sref[e.first], cnt);
            disp.release(x);
          });
        }
        for (auto& t : wgs) t.join();
      });
    }
    for (auto& t : xthreads) t.join();
    mem.kernel_end(cfg.late_wb);
    // ---- check (the next kernels read H): bit for bit, no NaN
    for (int i = 0; i < cfg.ninst; ++i) {
      Instance& I = *inst[i];
      for (int k = 0; k < I.nk; ++k)
        for (int j = 0; j < I.ni; ++j) {
          const int w = k / cfg.R, l = k % cfg.R;
          const uint32_t got = I.Hat(w, j + l, l).load();
          const float want = sref[i][(size_t)k * I.ni + j];
          if (std::isnan(fval(got))) cnt.nan++;
          else if (got != fbits(want)) cnt.wrong++;
        }
    }
    cnt.sweeps++;
  }
}

}  // namespace

int main(int argc, char** argv) {
  Cfg cfg;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto val = [&](const char* k) { return a.rfind(k, 0) == 0 ? std::atoi(a.c_str() + std::strlen(k)) : -1; };
    if (a == "--sentinel") cfg.tagged = false;
    else if (a == "--tagged") cfg.tagged = true;
    else if (a == "--late-writeback") cfg.late_wb = true;
    else if (a == "--plant-ring-bug") cfg.plant_ring_bug = true;
    else if (val("--sweeps=") >= 0) cfg.nsweep = val("--sweeps=");
    else if (val("--streams=") >= 0) cfg.streams = val("--streams=");
    else if (val("--slots=") >= 0) cfg.slots = val("--slots=");
    else if (val("--xcds=") >= 0) cfg.xcds = val("--xcds=");
    else if (val("--rows=") >= 0) cfg.R = val("--rows=");
    else if (val("--nk=") >= 0) cfg.nk = val("--nk=");
    else if (val("--ni=") >= 0) cfg.ni = val("--ni=");
    else if (val("--seed=") >= 0) cfg.seed = (unsigned)val("--seed=");
    else if (val("--wait-ms=") >= 0) cfg.wait_ms = val("--wait-ms=");
    else {
      std::fprintf(stderr, "unknown argument %s\n", a.c_str());
      return 2;
    }
  }
  if (cfg.R > 64 || cfg.R < 1 || cfg.S < 2 || cfg.nk < cfg.ninst + 1) return 2;
  Memory mem(cfg.xcds);
  if (cfg.late_wb) mem.start_writeback_agent();
  Dispatcher disp(cfg.xcds, cfg.slots);
  Counts cnt;
  std::atomic<int> linear_base{0};
  std::vector<std::thread> st;
  for (int s = 0; s < cfg.streams; ++s)
    st.emplace_back([&, s] { run_stream(cfg, s, mem, disp, linear_base, cnt); });
  for (auto& t : st) t.join();
  mem.stop_agent();
  std::printf("protocol=%s late_writeback=%d streams=%d sweeps=%ld wrong=%ld nan=%ld stale_accepts=%ld timeouts=%ld\n",
              cfg.tagged ? "tagged" : "sentinel", (int)cfg.late_wb, cfg.streams, cnt.sweeps.load(), cnt.wrong.load(),
              cnt.nan.load(), cnt.stale_accepts.load(), cnt.timeouts.load());
  const bool good = cnt.wrong == 0 && cnt.nan == 0 && cnt.timeouts == 0 && cnt.stale_accepts == 0;
  return good ? 0 : 1;
}
