// Batched stepping: several independent cases advance together, and every
// kernel launch of the step covers all the cases that issue it.
//
// Each case's host code (lw_step) runs unchanged on a thread of its own that
// records instead of launching (g_rec, lw_launch.h): a launch becomes a
// record of its kernel, grid, block and argument values; a zero fill or a
// device-to-host read joins the record; a host synchronisation is the
// batch's rendezvous. When every case has reached one, the last to arrive
// merges the recordings and issues them on the batch's stream:
// - the cases' sequences are merged greedily: of the operations at the head
//   of each case's sequence, the kernel most of them share is launched once
//   for all of them (k_bat), its workgroups cut into one range per case;
//   every case's own order is kept, so each case sees exactly the launches,
//   arguments and order it would see alone, and gives the same bits;
// - the argument records, the item tables and the by-reference argument
//   values travel in one host-to-device copy;
// - the read-backs are gathered on the device into one buffer and copied to
//   the host in one transfer, then handed to each case.
// The cases then continue to their next rendezvous. Cases whose host code
// decides differently (a multigrid batch more or less, a converged solve)
// simply record different sequences; the merge launches what they share
// together and the rest separately.
//
// Gates keep the cases in phase. The step passes a gate at the start of
// every substep and after the Poisson solve: a case that reaches a gate
// waits there until every case has reached one, while cases still waiting
// at a rendezvous go on (their next recordings, such as further multigrid
// cycles, run without the others). Each substep's work before and after the
// Poisson solve is therefore recorded by all cases between the same two
// gates and merged kernel by kernel.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lw_kernels.h"

void lw_set_last_error(const std::string &msg);

namespace lw {

thread_local Recorder *g_rec = nullptr;

// zero fill of n bytes (the step's hipMemsetAsync calls when recorded)
LW_KERNEL void k_zero_bytes(LW_KP(256, 0), char *p, long n) {
  const long nw = n >> 3;
  unsigned long long *w = (unsigned long long *)p;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nw; q += stride) w[q] = 0ull;
  for (long q = (nw << 3) + (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) p[q] = 0;
}

// a read-back: n bytes from the case's array into the batch's gather buffer
LW_KERNEL void k_copy_bytes(LW_KP(256, 0), char *dst, const char *src, long n) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
    dst[q] = src[q];
}

void rec_zero(void *p, size_t bytes) {
  if (bytes == 0) return;
  const long nw = (long)((bytes + 7) / 8);
  const unsigned grid = (unsigned)std::min<long>((nw + 255) / 256, 4096);
  rec_launch(LW_KFN(k_zero_bytes), "k_zero", dim3(grid), dim3(256), 0, (char *)p, (long)bytes);
}

void rec_read(void *host, const void *dev, size_t bytes) {
  if (bytes == 0) return;
  Recorder &r = *g_rec;
  RecOp op{};
  op.kind = 1;
  op.dst = host;
  op.src = dev;
  op.bytes = bytes;
  op.name = "read";
  r.ops.push_back(op);
}

struct Batch {
  int dev = 0;
  hipStream_t s = nullptr;
  int n = 0;
  int (*run)(void *, int, int64_t) = nullptr;
  void *ctx = nullptr;
  std::vector<Recorder> rec;
  std::vector<std::thread> th;
  std::mutex mu;
  std::condition_variable cv, cv_run, cv_done;
  int active = 0, arrived = 0, done = 0;
  std::vector<int> wait_kind;     // per case: 0 running, 1 at a rendezvous, 2 at a gate
  std::vector<uint64_t> release;  // per case: bumped when it may go on
  uint64_t gen = 0, run_gen = 0;
  int64_t nsteps = 0;
  bool quit = false;
  int err = 0;
  std::string msg;
  std::vector<int> rc;
  // staging: the flush's records, items and values (host pinned, device)
  char *hst = nullptr, *dst = nullptr;
  size_t cap = 0;
  char *hrb = nullptr, *drb = nullptr;  // read-back gather buffer
  size_t rcap = 0;
  // counters: flushes, batched launches, recorded launches, read-backs
  int64_t n_flush = 0, n_launch = 0, n_ops = 0, n_read = 0;
  int64_t us_prepare = 0, us_wait = 0;
};

thread_local Batch *g_batch = nullptr;
thread_local int g_member = -1;

namespace {

struct Key {
  int kind;
  const void *bat;
  unsigned bx, by, bz;
  bool operator==(const Key &o) const {
    return kind == o.kind && bat == o.bat && bx == o.bx && by == o.by && bz == o.bz;
  }
};
Key key_of(const RecOp &op) {
  if (op.kind == 1) return Key{1, nullptr, 0, 0, 0};
  return Key{0, op.bat, op.bx, op.by, op.bz};
}

struct Group {
  Key key;
  std::vector<std::pair<int, size_t>> ops;  // (recorder, op index)
};

size_t up16(size_t v) { return (v + 15) & ~(size_t)15; }

void grow(Batch *b, size_t need, char **h, char **d, size_t *cap) {
  if (need <= *cap) return;
  size_t c = std::max(need, *cap * 2);
  c = std::max<size_t>(c, 1 << 20);
  if (*h) (void)hipHostFree(*h);
  if (*d) (void)hipFree(*d);
  *h = nullptr;
  *d = nullptr;
  *cap = 0;
  if (hipHostMalloc((void **)h, c, hipHostMallocDefault) != hipSuccess || hipMalloc((void **)d, c) != hipSuccess)
    throw std::string("batch: staging allocation failed");
  *cap = c;
  (void)b;
}

// merges and issues every recorder's operations; called with mu held while
// every active member waits at its rendezvous
int64_t usec_since(std::chrono::steady_clock::time_point t0) {
  return (int64_t)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
}

void flush(Batch *b) {
  const auto t_start = std::chrono::steady_clock::now();
  const int nr = (int)b->rec.size();
  std::vector<size_t> pos(nr, 0);
  std::vector<Group> plan;
  for (;;) {
    // heads of the sequences and how many share each kernel
    std::vector<Key> keys;
    std::vector<int> cnt;
    std::vector<size_t> longest;
    for (int r = 0; r < nr; ++r) {
      if (pos[r] >= b->rec[r].ops.size()) continue;
      const Key k = key_of(b->rec[r].ops[pos[r]]);
      const size_t left = b->rec[r].ops.size() - pos[r];
      size_t q = 0;
      while (q < keys.size() && !(keys[q] == k)) ++q;
      if (q == keys.size()) {
        keys.push_back(k);
        cnt.push_back(0);
        longest.push_back(0);
      }
      cnt[q] += 1;
      longest[q] = std::max(longest[q], left);
    }
    if (keys.empty()) break;
    // the kernel most sequences start with; on a tie, the one whose
    // sequence has the most left (the case that lags goes first)
    size_t best = 0;
    for (size_t q = 1; q < keys.size(); ++q)
      if (cnt[q] > cnt[best] || (cnt[q] == cnt[best] && longest[q] > longest[best])) best = q;
    Group gp;
    gp.key = keys[best];
    for (int r = 0; r < nr; ++r) {
      if (pos[r] >= b->rec[r].ops.size()) continue;
      if (!(key_of(b->rec[r].ops[pos[r]]) == gp.key)) continue;
      gp.ops.emplace_back(r, pos[r]);
      pos[r] += 1;
    }
    plan.push_back(std::move(gp));
  }
  if (plan.empty()) return;

  // staging layout: [values of each recorder][per group: items, records][read-back records]
  std::vector<size_t> voff(nr, 0);
  size_t total = 0;
  for (int r = 0; r < nr; ++r) {
    voff[r] = total;
    total = up16(total + b->rec[r].vals.size());
  }
  size_t rbytes = 0;
  std::vector<size_t> goff(plan.size(), 0);
  for (size_t gi = 0; gi < plan.size(); ++gi) {
    const Group &gp = plan[gi];
    goff[gi] = total;
    total = up16(total + sizeof(BItem) * gp.ops.size());
    for (auto &ro : gp.ops) {
      const RecOp &op = b->rec[ro.first].ops[ro.second];
      if (op.kind == 1) {
        total = up16(total + 3 * sizeof(long));
        rbytes = up16(rbytes + op.bytes);
      } else {
        total = up16(total + op.arg_len);
      }
    }
  }
  grow(b, total, &b->hst, &b->dst, &b->cap);
  grow(b, std::max<size_t>(rbytes, 16), &b->hrb, &b->drb, &b->rcap);
  char *H = b->hst, *D = b->dst;
  for (int r = 0; r < nr; ++r)
    if (!b->rec[r].vals.empty()) memcpy(H + voff[r], b->rec[r].vals.data(), b->rec[r].vals.size());

  struct Launch {
    const void *fn;
    unsigned blocks;
    dim3 block;
    unsigned shm;
    const BItem *items;
    int nitems;
  };
  std::vector<Launch> launches;
  std::vector<std::pair<size_t, size_t>> scatter;  // (read-back offset, op) per read
  std::vector<std::pair<int, size_t>> rops;
  size_t roff = 0;
  const KFn<char *, const char *, long> kcopy = LW_KFN(k_copy_bytes);
  for (size_t gi = 0; gi < plan.size(); ++gi) {
    const Group &gp = plan[gi];
    size_t o = goff[gi];
    BItem *items = (BItem *)(H + o);
    size_t ro = up16(o + sizeof(BItem) * gp.ops.size());
    unsigned blk = 0, shm = 0;
    dim3 block(gp.key.bx, gp.key.by, gp.key.bz);
    for (size_t q = 0; q < gp.ops.size(); ++q) {
      const int r = gp.ops[q].first;
      const RecOp &op = b->rec[r].ops[gp.ops[q].second];
      BItem &it = items[q];
      it.blk0 = blk;
      if (op.kind == 1) {
        long *rec = (long *)(H + ro);
        rec[0] = (long)(b->drb + roff);
        rec[1] = (long)op.src;
        rec[2] = (long)op.bytes;
        it.gx = 1;
        it.gy = 1;
        it.gz = 1;
        it.args = D + ro;
        scatter.emplace_back(roff, rops.size());
        rops.emplace_back(r, gp.ops[q].second);
        roff = up16(roff + op.bytes);
        ro = up16(ro + 3 * sizeof(long));
      } else {
        const Recorder &rc = b->rec[r];
        memcpy(H + ro, rc.args.data() + op.arg_off, op.arg_len);
        for (uint32_t f = 0; f < op.fix_n; ++f) {
          const RecFix &fx = rc.fix[op.fix_off + f];
          const uint64_t addr = (uint64_t)(D + voff[r] + rc.val_off[fx.val]);
          memcpy(H + ro + (fx.slot - op.arg_off), &addr, 8);
        }
        it.gx = op.gx;
        it.gy = op.gy;
        it.gz = op.gz;
        it.args = D + ro;
        shm = std::max(shm, op.shm);
        ro = up16(ro + op.arg_len);
      }
      const uint64_t nb = (uint64_t)it.gx * it.gy * it.gz;
      const uint64_t next = (uint64_t)blk + ((nb + 7) & ~(uint64_t)7);
      if (next > 0x7fffffffull) throw std::string("batch: too many workgroups in one launch");
      blk = (unsigned)next;
    }
    if (gp.key.kind == 1) block = dim3(256);
    launches.push_back(Launch{gp.key.kind == 1 ? kcopy.bat_ : gp.key.bat, blk, block, shm,
                              (const BItem *)(D + o), (int)gp.ops.size()});
  }
  b->us_prepare += usec_since(t_start);
  const auto t_issue = std::chrono::steady_clock::now();
  hipError_t e = hipMemcpyAsync(D, H, total, hipMemcpyHostToDevice, b->s);
  for (const Launch &l : launches) {
    if (e != hipSuccess) break;
    const BItem *items = l.items;
    int nitems = l.nitems;
    void *args[2] = {(void *)&items, (void *)&nitems};
    e = hipLaunchKernel(l.fn, dim3(l.blocks), l.block, args, l.shm, b->s);
  }
  if (e == hipSuccess && roff > 0) e = hipMemcpyAsync(b->hrb, b->drb, roff, hipMemcpyDeviceToHost, b->s);
  if (e == hipSuccess) e = hipStreamSynchronize(b->s);
  b->us_wait += usec_since(t_issue);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) throw std::string("HIP error in a batched launch: ") + hipGetErrorString(e);
  for (size_t q = 0; q < scatter.size(); ++q) {
    const RecOp &op = b->rec[rops[q].first].ops[rops[q].second];
    memcpy(op.dst, b->hrb + scatter[q].first, op.bytes);
  }
  b->n_flush += 1;
  b->n_launch += (int64_t)launches.size();
  for (int r = 0; r < nr; ++r) {
    for (const RecOp &op : b->rec[r].ops) {
      if (op.kind == 1) b->n_read += 1;
      else b->n_ops += 1;
    }
    b->rec[r].clear();
  }
}

// with mu held, every active case waiting: runs what they recorded, then
// lets the cases at a rendezvous go on, or, when every case waits at a
// gate, all of them
void resolve_locked(Batch *b) {
  try {
    if (!b->err) flush(b);
  } catch (const std::string &m) {
    b->err = LW_ERR_DEVICE;
    b->msg = m;
  }
  if (b->err)
    for (auto &r : b->rec) r.clear();
  bool any_sync = false;
  for (int i = 0; i < b->n; ++i) any_sync = any_sync || b->wait_kind[i] == 1;
  for (int i = 0; i < b->n; ++i) {
    if (b->wait_kind[i] == 0) continue;
    if (b->err || !any_sync || b->wait_kind[i] == 1) {
      b->wait_kind[i] = 0;
      b->release[i] += 1;
      b->arrived -= 1;
    }
  }
  b->gen += 1;
  b->cv.notify_all();
}

// a recording case waits (kind 1 rendezvous, 2 gate) until released
void arrive(Batch *b, int i, int kind) {
  std::unique_lock<std::mutex> lk(b->mu);
  b->wait_kind[i] = kind;
  b->arrived += 1;
  const uint64_t r = b->release[i];
  if (b->arrived == b->active) resolve_locked(b);
  b->cv.wait(lk, [&] { return b->release[i] != r; });
  if (b->err) {
    lw_set_last_error(b->msg);
    throw (int)LW_ERR_DEVICE;
  }
}

void leave(Batch *b, int i) {
  std::unique_lock<std::mutex> lk(b->mu);
  b->rec[i].clear();
  b->active -= 1;
  if (b->active > 0 && b->arrived == b->active) resolve_locked(b);
}

void worker(Batch *b, int i) {
  (void)hipSetDevice(b->dev);
  uint64_t seen = 0;
  for (;;) {
    int64_t nsteps = 0;
    {
      std::unique_lock<std::mutex> lk(b->mu);
      b->cv_run.wait(lk, [&] { return b->quit || b->run_gen != seen; });
      if (b->quit) return;
      seen = b->run_gen;
      nsteps = b->nsteps;
    }
    g_batch = b;
    g_member = i;
    g_rec = &b->rec[i];
    const int rc = b->run(b->ctx, i, nsteps);
    g_rec = nullptr;
    g_batch = nullptr;
    leave(b, i);
    std::unique_lock<std::mutex> lk(b->mu);
    b->rc[i] = rc;
    b->done += 1;
    if (b->done == b->n) b->cv_done.notify_all();
  }
}

}  // namespace

void batch_sync() { arrive(g_batch, g_member, 1); }

void batch_gate() { arrive(g_batch, g_member, 2); }

Batch *batch_new(int dev, int n, int (*run)(void *, int, int64_t), void *ctx) {
  Batch *b = new Batch;
  b->dev = dev;
  b->n = n;
  b->run = run;
  b->ctx = ctx;
  b->rec.resize(n);
  b->rc.assign(n, 0);
  b->wait_kind.assign(n, 0);
  b->release.assign(n, 0);
  if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return nullptr;
  }
  for (int i = 0; i < n; ++i) b->th.emplace_back(worker, b, i);
  return b;
}

// every member runs its nsteps, recording; returns 0, or the first member's
// nonzero status in rc order
int batch_run(Batch *b, int64_t nsteps, int *rc) {
  {
    std::unique_lock<std::mutex> lk(b->mu);
    if (b->err) return b->err;
    b->nsteps = nsteps;
    b->active = b->n;
    b->arrived = 0;
    std::fill(b->wait_kind.begin(), b->wait_kind.end(), 0);
    b->done = 0;
    b->run_gen += 1;
    b->cv_run.notify_all();
    b->cv_done.wait(lk, [&] { return b->done == b->n; });
  }
  int first = 0;
  for (int i = 0; i < b->n; ++i) {
    if (rc) rc[i] = b->rc[i];
    if (!first && b->rc[i]) first = b->rc[i];
  }
  return first;
}

const std::string &batch_error(const Batch *b) { return b->msg; }

void batch_counts(const Batch *b, int64_t out[6]) {
  out[0] = b->n_flush;
  out[1] = b->n_launch;
  out[2] = b->n_ops;
  out[3] = b->n_read;
  out[4] = b->us_prepare;
  out[5] = b->us_wait;
}

void batch_free(Batch *b) {
  if (!b) return;
  {
    std::unique_lock<std::mutex> lk(b->mu);
    b->quit = true;
    b->cv_run.notify_all();
  }
  for (auto &t : b->th) t.join();
  if (b->s) (void)hipStreamSynchronize(b->s);
  if (b->s) (void)hipStreamDestroy(b->s);
  if (b->hst) (void)hipHostFree(b->hst);
  if (b->dst) (void)hipFree(b->dst);
  if (b->hrb) (void)hipHostFree(b->hrb);
  if (b->drb) (void)hipFree(b->drb);
  delete b;
}

}  // namespace lw
