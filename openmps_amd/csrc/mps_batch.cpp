// Several independent scenes stepped together (mps_batch_*): Computer::
// ForwardTime (Computer.hpp:1700-1751) of every scene, with the pressure
// solves of all scenes on the mixed multilevel solver as one set of grouped
// launches (mps_cg.hip: batch_solve) and one MLP inference over the feature
// rows of all scenes on the learned warm start.
//
// A step is three phases:
//   1. per scene, on its own stream, from a pool of host threads: the CFL step
//      (DetermineDt, :759-777) or the fixed one, the neighbour search, both
//      density passes, the explicit stage, the pressure system, the learned
//      start's features and the multilevel set-up (or, for a scene on another
//      solver, its own solve);
//   2. on the batch's stream: the inference over all scenes' rows and the
//      batched solve (each grouped launch covers every scene still iterating);
//   3. per scene again: pressure write-back, pressure gradient and DS.
// The scenes keep their own device buffers; the batch only adds its argument
// tables, statuses and the shared feature rows.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <exception>
#include <functional>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "mps_internal.hpp"

namespace mps {
namespace {

// Host threads that run one job over a list of scene indices (largest
// scenes first, taken dynamically) and return when all are done.
class Pool {
 public:
  Pool(int threads, int device) : device_(device) {
    for (int k = 0; k < threads; k++) th_.emplace_back([this] { loop(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  void run(const std::vector<size_t>& order, const std::function<void(size_t)>& job) {
    if (th_.empty()) {
      for (size_t k : order) job(k);
      return;
    }
    std::unique_lock<std::mutex> lk(mu_);
    order_ = &order;
    job_ = &job;
    next_ = 0;
    pending_ = order.size();
    gen_++;
    cv_.notify_all();
    done_cv_.wait(lk, [this] { return pending_ == 0; });
    job_ = nullptr;
  }

 private:
  void loop() {
    (void)hipSetDevice(device_);
    long long seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      while (job_ && next_ < order_->size()) {
        const size_t k = (*order_)[next_++];
        const std::function<void(size_t)>* job = job_;
        lk.unlock();
        (*job)(k);
        lk.lock();
        if (--pending_ == 0) done_cv_.notify_all();
      }
    }
  }
  int device_;
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::vector<size_t>* order_ = nullptr;
  const std::function<void(size_t)>* job_ = nullptr;
  size_t next_ = 0, pending_ = 0;
  long long gen_ = 0;
  bool stop_ = false;
};

}  // namespace
}  // namespace mps

struct mps_batch {
  int device = 0;
  std::vector<mps_ctx*> scenes;
  std::vector<hipEvent_t> ev;  // per scene: its phase-1 work
  hipEvent_t ev_solved = nullptr;
  mps::BatchWork w;
  mps::Pool* pool = nullptr;
  // the learned warm start's feature rows and predictions of all scenes
  __bf16* feats = nullptr;
  float* ypred = nullptr;
  size_t feats_cap = 0;
  std::string err;
};

namespace mps {
namespace {

struct SceneError {
  int code = MPS_OK;
  std::string msg;
};

bool learned_scene(const Context& c) {
  return c.solver.warm_start == MPS_WARM_LEARNED && c.model && c.solver.kind == MPS_SOLVER_MIXED_IR &&
         c.solver.precond == MPS_PRECOND_MULTILEVEL;
}

void give_back_features(mps_batch& b) {
  for (mps_ctx* h : b.scenes) {
    Context& c = h->c;
    if (!c.feats_borrowed) continue;
    c.feats = nullptr;
    c.ypred = nullptr;
    c.feats_cap = 0;
    dfree(c.feat_rec);
    c.feats_borrowed = false;
  }
}

void batch_forward(mps_batch& b, double dt_fixed, double* dt_out) {
  const size_t S = b.scenes.size();
  if (S == 0) return;
  MPS_HIP(hipSetDevice(b.device));
  std::vector<size_t> order(S);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t c) { return b.scenes[a]->c.n > b.scenes[c]->c.n; });
  // rows of the shared feature buffer: the learned scenes' particles, in scene order
  std::vector<size_t> row0(S, 0);
  size_t rows = 0;
  Model* model = nullptr;
  for (size_t s = 0; s < S; s++) {
    Context& c = b.scenes[s]->c;
    if (!learned_scene(c)) continue;
    if (model && c.model != model)
      throw Failure{MPS_ERR_BAD_ARG, "scenes of a batch on the learned start must share one model"};
    model = c.model;
    row0[s] = rows;
    rows += c.n;
  }
  if (rows > b.feats_cap) {
    give_back_features(b);
    dfree(b.feats);
    dfree(b.ypred);
    b.feats = dalloc<__bf16>(rows * kFeatPad);
    b.ypred = dalloc<float>(rows);
    b.feats_cap = rows;
  }
  std::vector<SceneError> err(S);
  std::vector<char> joined(S, 0), learned(S, 0);
  // 1. per scene: the stages before the solve
  b.pool->run(order, [&](size_t s) {
    Context& c = b.scenes[s]->c;
    try {
      const double dt = (dt_fixed > 0) ? dt_fixed : max_speed_dt(c);
      if (dt_out) dt_out[s] = dt;
      c.dt = dt;
      c.t += dt;
      search(c);
      density_and_explicit(c);
      density_save_and_assemble(c);
      if (batch_eligible(c)) {
        c.stats = mps_solve_stats{};
        if (learned_scene(c)) {
          c.feats = b.feats + row0[s] * kFeatPad;
          c.ypred = b.ypred + row0[s];
          c.feats_borrowed = true;
          learned_features(c);
          learned[s] = 1;
        }
        if (batch_amg(c)) amg_prepare(c);
        else ml_prepare(c);
        MPS_HIP(hipEventRecord(b.ev[s], c.stream));
        joined[s] = 1;
      } else {
        ppe_solve(c);
      }
    } catch (Failure& f) {
      err[s] = SceneError{f.code, f.msg};
    } catch (std::exception& e) {
      err[s] = SceneError{MPS_ERR_BAD_ARG, e.what()};
    }
  });
  // 2. the batch: one inference over the learned scenes' rows, then the solve
  std::vector<Context*> solve, lrn;
  for (size_t s = 0; s < S; s++) {
    if (!joined[s] || err[s].code != MPS_OK) continue;
    MPS_HIP(hipStreamWaitEvent(b.w.stream, b.ev[s], 0));
    solve.push_back(&b.scenes[s]->c);
    if (learned[s]) lrn.push_back(&b.scenes[s]->c);
  }
  int code = MPS_OK;
  std::string msg;
  try {
    if (!lrn.empty()) {
      // the rows of scenes that failed above are computed but not used
      batch_learned_start(lrn, *model, b.w.stream, b.feats, b.ypred, rows);
    }
    batch_solve(solve, b.w);  // synchronises the batch's stream
    for (Context* c : lrn) batch_learned_outcome(*c);
  } catch (Failure& f) {
    code = f.code;
    msg = f.msg;
    if (code != MPS_ERR_CG_NOT_CONVERGED) throw;
  }
  MPS_HIP(hipEventRecord(b.ev_solved, b.w.stream));
  // 3. per scene: the stages after the solve
  b.pool->run(order, [&](size_t s) {
    Context& c = b.scenes[s]->c;
    if (err[s].code != MPS_OK) return;
    try {
      if (joined[s]) MPS_HIP(hipStreamWaitEvent(c.stream, b.ev_solved, 0));
      pressure_writeback(c);
      pressure_gradient(c);
      dynamic_stabilize(c);
      MPS_HIP(hipStreamSynchronize(c.stream));
      collect_events(c);
    } catch (Failure& f) {
      err[s] = SceneError{f.code, f.msg};
    } catch (std::exception& e) {
      err[s] = SceneError{MPS_ERR_BAD_ARG, e.what()};
    }
  });
  for (size_t s = 0; s < S; s++)
    if (err[s].code != MPS_OK) {
      char head[64];
      std::snprintf(head, sizeof head, "scene %zu: ", s);
      throw Failure{err[s].code, head + err[s].msg};
    }
  if (code != MPS_OK) throw Failure{code, msg};
}

}  // namespace
}  // namespace mps

using mps::Failure;

extern "C" {

int mps_batch_create(int device, int host_threads, mps_batch** out) {
  if (!out) return MPS_ERR_BAD_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return MPS_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= ndev || host_threads < 0 || host_threads > 256) return MPS_ERR_BAD_ARG;
  mps_batch* b = new (std::nothrow) mps_batch;
  if (!b) return MPS_ERR_BAD_ARG;
  try {
    MPS_HIP(hipSetDevice(device));
    b->device = device;
    b->w.device = device;
    b->w.ensure(8);
    MPS_HIP(hipEventCreateWithFlags(&b->ev_solved, hipEventDisableTiming));
    b->pool = new mps::Pool(host_threads, device);
  } catch (Failure& f) {
    std::fprintf(stderr, "mps_batch_create: %s\n", f.msg.c_str());
    b->w.release();
    if (b->ev_solved) (void)hipEventDestroy(b->ev_solved);
    delete b;
    return f.code;
  }
  *out = b;
  return MPS_OK;
}

void mps_batch_destroy(mps_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  delete b->pool;
  if (b->w.stream) (void)hipStreamSynchronize(b->w.stream);
  mps::give_back_features(*b);
  mps::dfree(b->feats);
  mps::dfree(b->ypred);
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  if (b->ev_solved) (void)hipEventDestroy(b->ev_solved);
  b->w.release();
  delete b;
}

const char* mps_batch_last_error(const mps_batch* b) { return b ? b->err.c_str() : "null batch"; }

int mps_batch_add(mps_batch* b, mps_ctx* scene) {
  if (!b || !scene) return MPS_ERR_BAD_ARG;
  try {
    if (scene->c.device != b->device) throw Failure{MPS_ERR_BAD_ARG, "the scene's context is on another device"};
    if (std::find(b->scenes.begin(), b->scenes.end(), scene) != b->scenes.end())
      throw Failure{MPS_ERR_BAD_ARG, "the scene is in the batch already"};
    MPS_HIP(hipSetDevice(b->device));
    hipEvent_t e = nullptr;
    MPS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->ev.push_back(e);
    b->scenes.push_back(scene);
    b->w.ensure(b->scenes.size());
    b->err.clear();
    return MPS_OK;
  } catch (Failure& f) {
    b->err = f.msg;
    return f.code;
  }
}

int mps_batch_forward_time(mps_batch* b, double dt, double* dt_out) {
  if (!b) return MPS_ERR_BAD_ARG;
  try {
    mps::batch_forward(*b, dt, dt_out);
    b->err.clear();
    return MPS_OK;
  } catch (Failure& f) {
    b->err = f.msg;
    if (f.code == MPS_ERR_HIP) (void)hipGetLastError();
    return f.code;
  } catch (std::exception& e) {
    b->err = e.what();
    return MPS_ERR_BAD_ARG;
  }
}

int mps_batch_solve_info(const mps_batch* b, double* ms, int64_t* launches) {
  if (!b) return MPS_ERR_BAD_ARG;
  if (ms) *ms = b->w.solve_ms;
  if (launches) *launches = b->w.launches;
  return MPS_OK;
}

}  // extern "C"
