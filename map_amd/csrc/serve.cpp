// Serving loop: map-algebra requests evaluated as they arrive.
//
// Each request is planned when admitted (fusion groups, levels, kernels) and
// admitted while its HBM footprint (the rasters its kernels write, inputs to
// upload, zonal partials) fits the budget and the in-flight limit allows.
// Every step enqueues the next level of every admitted request on the one
// compute stream, batching the groups of all requests by kernel into single
// launches over a device task table (continuous batching: a request admitted
// while others run joins the next launches), records an event per step, and
// retires a request once the stream has passed its last step: zonal results
// are handed out then, its intermediates released, its time recorded.
// Requests share resident rasters (cached DEMs, results of earlier requests
// that are still referenced); a request that needs a node another admitted
// request is still computing waits for that request.
#include "dev.hpp"
#include "runtime.hpp"
#include <chrono>
#include <deque>
#include <sstream>
#include <algorithm>
#include <stdexcept>
#include <thread>

namespace mapx {

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Runtime::Request {
  long long id = 0;
  Eval ev;
  size_t next = 0;         // next level to enqueue
  size_t footprint = 0;    // bytes counted against the budget
  long last_step = -1;     // step of its last enqueued work
  double t_submit = 0, t_admit = 0;
  std::vector<Node*> computes;  // nodes this request computes
  std::vector<Node*> pinned;    // resident rasters it reads
  Request* blocker = nullptr;   // waiting: the admitted request computing a node it needs
  unsigned long long tick = 0;  // while admitted: its nodes stay in HBM (Runtime::spill_one)
};

struct Runtime::ServeState {
  size_t budget = 0, inflight = 0;
  int max_inflight = 0;
  long long next_id = 0;
  std::deque<std::unique_ptr<Request>> queued;  // in arrival (id) order
  std::vector<std::unique_ptr<Request>> active;
  // requests that need a node an admitted request computes, by that request;
  // they return to `queued` when it retires (so admission never rescans them)
  std::unordered_map<Request*, std::vector<std::unique_ptr<Request>>> waiting;
  struct Step {
    void* ev = nullptr;
    long id = 0;
    std::vector<Readback> rbs;
  };
  std::deque<Step> steps;  // enqueued steps the stream has not passed yet
  long step_no = 0, step_done = -1;
  std::unordered_map<Node*, Request*> owner;
  std::vector<ServeDone> done;
  long admitted = 0, completed = 0, launch_steps = 0, tasks = 0, waits_memory = 0, waits_owner = 0;
  long done_held = 0;  // requests whose targets were all held at admission (no work)
  size_t peak_inflight = 0;
};

Runtime::Runtime() = default;
Runtime::~Runtime() = default;

void Runtime::serveOpen(size_t hbm_budget, int max_inflight) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  if (!dev::initialized()) throw std::runtime_error("serveOpen: devices not set up");
  if (srv_) throw std::runtime_error("serveOpen: a server is already open");
  if (max_inflight < 1) throw std::runtime_error("serveOpen: max_inflight must be >= 1");
  srv_ = std::make_unique<ServeState>();
  srv_->budget = hbm_budget ? hbm_budget : (size_t)(0.8 * (double)dev::mem_free());
  srv_->max_inflight = max_inflight;
}

long long Runtime::serveSubmit(const std::vector<Node*>& targets) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  if (!srv_) throw std::runtime_error("serveSubmit: no server is open");
  if (targets.empty()) throw std::runtime_error("serveSubmit: nothing to evaluate");
  auto r = std::make_unique<Request>();
  r->id = srv_->next_id++;
  r->ev.targets = with_aux(targets);
  for (Node* t : r->ev.targets) t->ref++;  // held until the request retires (released in serve_retire)
  r->t_submit = now_s();
  srv_->queued.push_back(std::move(r));
  return srv_->next_id - 1;
}

void Runtime::serve_retire() {
  ServeState& S = *srv_;
  while (!S.steps.empty() && dev::event_done(S.steps.front().ev)) {
    for (Readback& rb : S.steps.front().rbs) distribute(rb);
    dev::event_destroy(S.steps.front().ev);
    S.step_done = S.steps.front().id;
    S.steps.pop_front();
  }
  const double t = now_s();
  for (size_t i = 0; i < S.active.size();) {
    Request& r = *S.active[i];
    if (r.next < r.ev.levels.size() || r.last_step > S.step_done) { i++; continue; }
    eval_finish(r.ev);
    for (Node* n : r.computes) {
      auto it = S.owner.find(n);
      if (it != S.owner.end() && it->second == &r) S.owner.erase(it);
    }
    for (Node* n : r.pinned)
      if (--pins_[n] == 0) pins_.erase(n);
    tick_end(r.tick);
    for (Node* n : r.ev.targets) n->ref--;
    S.inflight -= r.footprint;
    S.done.push_back({r.id, r.t_submit, r.t_admit, t});
    S.completed++;
    auto w = S.waiting.find(&r);
    if (w != S.waiting.end()) {  // back into the queue at their arrival positions
      for (auto& q : w->second) {
        q->blocker = nullptr;
        auto pos = std::lower_bound(S.queued.begin(), S.queued.end(), q->id,
                                    [](const std::unique_ptr<Request>& a, long long id) { return a->id < id; });
        S.queued.insert(pos, std::move(q));
      }
      S.waiting.erase(w);
    }
    S.active[i] = std::move(S.active.back());
    S.active.pop_back();
  }
}

int Runtime::serveStep(int wait_us) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  if (!srv_) throw std::runtime_error("serveStep: no server is open");
  ServeState& S = *srv_;
  const long done0 = S.completed;
  serve_retire();
  release_unreferenced();
  dag.collect_grown([this](Node* n) { forget(n); });

  // admission, in arrival order (a request blocked by another's node lets later ones pass)
  for (auto it = S.queued.begin(); it != S.queued.end();) {
    if ((int)S.active.size() >= S.max_inflight) break;
    Request& r = **it;
    {
      // targets already held (computed by an earlier request: value numbering
      // gives identical requests the same nodes) need no work; their hold is
      // released now, and a request left with none is done without planning
      std::vector<Node*> left;
      for (Node* n : r.ev.targets) {
        const bool held = (n->meta.nd == D0 && n->evaluated && n->kind != Kind::Constant) ||
                          (n->meta.nd == D2 && n->kind != Kind::Read && (resident_.count(n) || spilled_.count(n)));
        if (held) n->ref--;
        else left.push_back(n);
      }
      r.ev.targets = std::move(left);
      if (r.ev.targets.empty()) {
        const double t = now_s();
        S.done.push_back({r.id, r.t_submit, t, t});
        S.completed++;
        S.admitted++;
        S.done_held++;
        it = S.queued.erase(it);
        continue;
      }
    }
    Eval e;
    e.targets = r.ev.targets;
    eval_plan(e);
    for (Node* n : e.list) {
      if (n->cut) continue;
      auto ow = S.owner.find(n);
      if (ow != S.owner.end()) { r.blocker = ow->second; break; }
    }
    if (r.blocker) {  // planned again once that request has retired
      eval_uncut(e);
      S.waits_owner++;
      Request* b = r.blocker;
      S.waiting[b].push_back(std::move(*it));
      it = S.queued.erase(it);
      continue;
    }
    size_t fp = 0;
    std::vector<Node*> pins;
    // marked before any input is brought back from host memory, so that doing
    // so never moves out another raster this request reads
    const unsigned long long tick = tick_begin(e.list);
    try {
      for (auto& g : e.plan->groups) {
        if (host_group(*g)) continue;
        Node* rn = radial_node(*g);
        Node* sn = spread_node(*g);
        if (rn || sn) {  // no generated kernel: its output raster (and a spread's count words)
          bool z = false;
          const Meta& m = (rn ? rn : sn)->meta;
          if (!kernel_outputs(*g, e.tsorted, z).empty())
            fp += rasterBytes(m) + (sn ? (size_t)m.bs.arr[0] * m.bs.arr[1] * m.nb(0) * m.nb(1) * 8 + 8 : 0);
          for (Node* n : g->ins) {
            if (n->meta.nd == D2 && spilled_.count(n)) unspill(n);
            if (n->meta.nd == D2 && resident_.count(n)) pins.push_back(n);
            else if (n->meta.nd == D2 && n->kind == Kind::Read) fp += rasterBytes(n->meta);
          }
          continue;
        }
        bool zonal = false;
        std::vector<Node*> kouts = kernel_outputs(*g, e.tsorted, zonal);
        KernelEntry* ke = nullptr;
        if (!kouts.empty() || zonal) ke = &kernel_for(e.plan.get(), *g, kouts);
        e.kern[g.get()] = {kouts, ke};
        for (Node* n : kouts) fp += rasterBytes(n->meta);
        if (zonal) fp += (size_t)(rasterBytes(g->meta()) / 4096 + 1) * ke->spec.part_bytes;
        for (Node* n : g->ins) {
          if (n->meta.nd != D2) continue;
          if (spilled_.count(n)) unspill(n);                       // back in HBM for the request
          if (resident_.count(n)) pins.push_back(n);               // read, never dropped while admitted
          else if (n->kind == Kind::Read) fp += rasterBytes(n->meta);  // uploaded when first read
        }
      }
    } catch (...) {
      tick_end(tick);
      eval_uncut(e);
      throw;
    }
    eval_uncut(e);
    if (!S.active.empty() && S.inflight + fp > S.budget) {
      tick_end(tick);
      S.waits_memory++;
      break;  // memory frees up as requests retire
    }
    r.tick = tick;
    r.ev = std::move(e);
    r.footprint = fp;
    r.t_admit = now_s();
    for (auto& g : r.ev.plan->groups)
      for (Node* n : g->nodes) {
        r.computes.push_back(n);
        S.owner[n] = &r;
      }
    std::sort(pins.begin(), pins.end());
    pins.erase(std::unique(pins.begin(), pins.end()), pins.end());
    for (Node* n : pins) pins_[n]++;
    r.pinned = pins;
    S.inflight += fp;
    S.peak_inflight = std::max(S.peak_inflight, S.inflight);
    S.admitted++;
    S.active.push_back(std::move(*it));
    it = S.queued.erase(it);
  }

  // enqueue: levels of all admitted requests whose inputs are on hand, one
  // step per round, until no request can advance without the host
  for (;;) {
    std::vector<std::pair<Eval*, size_t>> work;
    std::vector<Request*> who;
    for (auto& rp : S.active) {
      Request& r = *rp;
      if (r.next >= r.ev.levels.size()) continue;
      bool ready = true;
      for (const Group* g : r.ev.levels[r.next])
        for (Node* n : g->ins)
          if (n->meta.nd == D0 && !n->evaluated) ready = false;
      if (!ready) continue;  // waits for zonal results of an earlier step
      work.push_back({&r.ev, r.next});
      who.push_back(&r);
    }
    if (work.empty()) break;
    ServeState::Step st;
    st.id = S.step_no++;
    for (auto& w : work) S.tasks += (long)w.first->levels[w.second].size();
    run_levels(work, &st.rbs);
    for (size_t i = 0; i < work.size(); i++) {
      eval_release(*work[i].first, work[i].second);
      who[i]->next++;
      who[i]->last_step = st.id;
    }
    st.ev = dev::event_record();
    S.steps.push_back(std::move(st));
    S.launch_steps++;
  }

  // completion detection
  if (wait_us > 0 && !S.steps.empty()) {
    const double deadline = now_s() + wait_us * 1e-6;
    while (!dev::event_done(S.steps.front().ev) && now_s() < deadline)
      std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  serve_retire();
  return (int)(S.completed - done0);
}

std::vector<Runtime::ServeDone> Runtime::serveTake(size_t cap) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  if (!srv_) throw std::runtime_error("serveTake: no server is open");
  auto& d = srv_->done;
  const size_t n = std::min(cap, d.size());
  std::vector<ServeDone> out(d.begin(), d.begin() + n);
  d.erase(d.begin(), d.begin() + n);
  return out;
}

std::string Runtime::serveStats() {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  if (!srv_) throw std::runtime_error("serveStats: no server is open");
  const ServeState& S = *srv_;
  size_t nwait = 0;
  for (const auto& w : S.waiting) nwait += w.second.size();
  std::ostringstream s;
  s << "{\"queued\": " << S.queued.size() << ", \"waiting\": " << nwait << ", \"active\": " << S.active.size() << ", \"admitted\": " << S.admitted
    << ", \"completed\": " << S.completed << ", \"steps\": " << S.launch_steps << ", \"group_tasks\": " << S.tasks
    << ", \"inflight_bytes\": " << S.inflight << ", \"peak_inflight_bytes\": " << S.peak_inflight
    << ", \"budget_bytes\": " << S.budget << ", \"waits_memory\": " << S.waits_memory
    << ", \"waits_owner\": " << S.waits_owner << ", \"done_held\": " << S.done_held << "}";
  return s.str();
}

void Runtime::serveClose() {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  if (!srv_) return;
  // drain: every submitted request runs to completion
  while (!srv_->queued.empty() || !srv_->active.empty()) {
    const long before = srv_->completed + srv_->admitted + srv_->step_no;
    serveStep(2000);
    if (srv_->steps.empty() && srv_->completed + srv_->admitted + srv_->step_no == before)
      throw std::runtime_error("serveClose: requests cannot make progress");
  }
  for (auto& st : srv_->steps) dev::event_destroy(st.ev);
  srv_.reset();
}

}  // namespace mapx
