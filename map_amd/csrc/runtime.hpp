// Block-task execution runtime: evaluation driver, device-resident raster
// store, per-group kernel launches (plain / batched / graph forms), host
// scalar tasks and zonal finalisation.
//
// Replaces the reference's Runtime::evaluate/workflow (runtime/Runtime.cpp:397-507),
// Program/Version (kernel generation and compile cache), Scheduler/Worker
// (job loop) and Cache/Block (device block cache). Rasters stay in HBM whole
// and block-tiled between fused groups; no host round trip per block.
#pragma once
#include "codegen.hpp"
#include "dag.hpp"
#include <functional>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace mapx {

enum ExecForm : int {
  FORM_PLAIN = 0,    // one launch per fused group per block (the reference's job granularity)
  FORM_BATCHED = 1,  // one launch per fused group covering all its blocks
};

struct Options {
  int form = FORM_BATCHED;
  bool graph = false;         // replay identical launch sequences from hipGraphs
  bool focal_zonal = false;   // fuse FOCAL -> ZONAL into one kernel
  bool code_fusion = true;
  int wg_per_cu = 0;          // persistent grid size in work-groups per CU (0: the kernel's occupancy)
  bool batch_tasks = true;    // independent groups of one kernel share one launch (task table)
  int queue = 0;              // device chunk queue: 0 for batched launches only, 1 always, 2 never
  int chunk_tiles = 0;        // tiles per chunk of non-zonal groups (0: 16)
  TileShape tile{0, 0, 0, 0, 0}; // forced tile shape (TX 0 = per-group default / tuned)
  bool libm = false;          // FP32 sqrt / atan2 / sin-cos pairs through the device library, not the fast paths
  bool predict = true;        // zonal reductions of constant rasters computed on the host (no kernel)
  size_t hbm_cap = 0;         // bytes of HBM the raster store may hold (0: as much as the device gives)
};

struct Counters {
  long evals = 0, groups = 0, launches = 0, scalar_tasks = 0, compiles = 0;
  long shape_hits = 0;  // kernels taken from the shape cache instead of generated
  long predicted = 0;   // zonal reductions of constant rasters known without a kernel
  long narrow_launches = 0;  // batches run by the module for edge blocks narrower than the halo
  long bytes_h2d = 0, bytes_d2h = 0;
  long spills = 0, unspills = 0, dropped_for_space = 0;  // resident rasters moved out of HBM and back
  long bytes_spilled = 0;
  double ms_last_eval = 0;
  // host time per phase (the reference's Clock regions, runtime/Clock.hpp:24-27)
  double ms_fuse = 0, ms_codegen = 0, ms_launch = 0, ms_eval_host = 0;
  double ms_plan = 0, ms_prepare = 0;  // evaluation planning (cut, plan lookup, levels); task preparation
  double ms_collect = 0, ms_predict = 0;  // releases and DAG collection before an evaluation; aux targets + prediction
};

// Stream-ordered caching allocator: all device work runs on one in-order
// stream, so a released buffer can be handed to later work at once. When the
// device (or the cap) has no room, cached buffers go back to the device first,
// then `reclaim` is asked to move a resident raster out of HBM, until the
// allocation fits or nothing is left to move.
class Store {
 public:
  void* alloc(size_t bytes);
  void release(void* p);
  void discard(void* p);  // a buffer in use goes back to the device now
  void trim();        // returns cached buffers to the device
  size_t in_use() const { return in_use_; }
  size_t cached() const { return cached_; }
  size_t cap = 0;     // bytes held (in use + cached) at most; 0: no cap
  std::function<bool()> reclaim;  // frees some store memory; false when it cannot
  ~Store();
 private:
  std::multimap<size_t, void*> free_;
  std::unordered_map<void*, size_t> size_;
  size_t in_use_ = 0, cached_ = 0;
};

class Runtime {
 public:
  static Runtime& get();
  Runtime();
  ~Runtime();

  Dag dag;
  Options opt;
  Counters cnt;
  Store store;

  void setupDevices(const std::string& plat, unsigned dev_type, const std::string& dev_name);
  void evaluate(const std::vector<Node*>& targets);

  // D2 node whose raster is held after evaluation, while referenced: in HBM or, moved out for room, in host memory
  bool isResident(const Node* n) const {
    Node* m = const_cast<Node*>(n);
    return resident_.count(m) != 0 || spilled_.count(m) != 0;
  }
  void* residentPtr(Node* n);
  size_t rasterBytes(const Meta& m) const;
  static Node* radial_node(const Group& g);  // the RadialScan node of a radial group, or null
  static Node* spread_node(const Group& g);  // the SpreadScan node of a spreading group, or null
  static bool host_group(const Group& g);    // evaluated by scalar_task (D0 values, cell accesses)
  void fetch(Node* n, void* host, size_t bytes);  // copies a resident raster (block-tiled) to host
  void drop(Node* n);                              // releases a node's resident raster
  void adopt(Node* n, void* dev) { resident_[n] = dev; }  // raster produced in HBM by a generator
  void release_unreferenced();
  void forget(Node* n);  // a node is being destroyed: its raster goes back to the store

  // Plan of an evaluation without executing it (groups and kernel sources), as JSON.
  std::string plan(const std::vector<Node*>& targets);

  // Autotuner: evaluates `targets` once per candidate tile shape (TX, TY, VX),
  // times every launch with HIP events and keeps, per fused-group signature
  // and block size, the fastest shape; tile selection consults the result.
  // Returns a JSON report of every candidate's average kernel time.
  std::string autotune(const std::vector<Node*>& targets, int reps);
  void tune_load(const std::string& path);
  void tune_save(const std::string& path) const;
  size_t tuned_count() const { return tuned_.size(); }

  // ---- serving: requests evaluated as they arrive, batched across requests ----
  // A request is a set of target nodes (a script over resident rasters).
  // Requests are admitted while their HBM footprint fits the budget and fewer
  // than max_inflight run; every step enqueues the next level of every
  // admitted request, batching the groups of all requests by kernel, and
  // retires requests whose work the stream has passed. No step waits for the
  // GPU except to detect completions (up to wait_us).
  struct ServeDone {
    long long id;
    double t_submit, t_admit, t_done;  // steady-clock seconds (CLOCK_MONOTONIC)
  };
  void serveOpen(size_t hbm_budget, int max_inflight);
  long long serveSubmit(const std::vector<Node*>& targets);
  int serveStep(int wait_us);
  std::vector<ServeDone> serveTake(size_t cap);  // up to cap retired requests, oldest first
  std::string serveStats();
  void serveClose();
  bool serving() const { return srv_ != nullptr; }

  // ---- resident cache: rasters kept in HBM under caller keys ----
  // (DEMs shared by the requests of a server, or by successive evaluations):
  // the cache holds a reference on each node; past the byte budget (0: no
  // limit) the least recently used entries are released, and their rasters
  // leave HBM once no request or caller holds them.
  void cacheBudget(size_t bytes);
  Node* cacheGet(long long key);      // null on a miss; a hit becomes the most recent entry
  void cachePut(long long key, Node* n);
  void cacheDrop(long long key);
  std::string cacheStats() const;
  std::string kernelSource(const std::string& name) const {
    auto it = kernel_src_.find(name);
    return it == kernel_src_.end() ? std::string() : it->second;
  }

 private:
  struct Task;
  // Zonal results of one batch on their way to the host.
  struct Readback {
    unsigned char* host = nullptr;  // pinned, one record of PB bytes per task
    size_t PB = 0;
    std::vector<std::vector<ZonalOut>> outs;  // per task: its zonal output nodes and record offsets
  };
  struct Plan;
  struct KernelEntry;
  // One evaluation in progress: its plan, levels and the use counts of its rasters.
  struct Eval {
    std::vector<Node*> targets, list;
    std::vector<Node*> tsorted;  // the targets sorted by address (d2_outputs)
    std::shared_ptr<Plan> plan;
    std::vector<std::vector<const Group*>> levels;
    std::unordered_map<Node*, int> uses;
    std::unordered_set<Node*> tset;
    // kernels fixed at planning time (requests of the server run later)
    std::unordered_map<const Group*, std::pair<std::vector<Node*>, KernelEntry*>> kern;
  };
  std::vector<Node*> cut_list(const std::vector<Node*>& targets);
  void eval_plan(Eval& e);
  void eval_uncut(Eval& e);
  void eval_release(Eval& e, size_t level);
  void eval_finish(Eval& e);
  std::vector<Node*> kernel_outputs(const Group& g, const std::vector<Node*>& targets, bool& has_zonal);
  void run_levels(const std::vector<std::pair<Eval*, size_t>>& work, std::vector<Readback>* async);
  static constexpr int kBigLevel = 64;    // levels of at least this many groups run unrecorded
  static constexpr int kSubBatch = 512;   // groups prepared and launched together when unrecorded
  static constexpr int kFirstSubBatch = 64;  // the first of them
  bool prepare_task(const Group& g, Eval& e, Task& t);
  void radial_task(const Group& g, Eval& e);
  void spread_task(const Group& g, Eval& e);
  std::vector<Node*> with_aux(const std::vector<Node*>& targets);
  // Constant-raster prediction: zonal reductions of rasters that hold one
  // value everywhere get their value without a kernel; returns their count.
  int predict(const std::vector<Node*>& targets);
  Variant access_value(Node* n);
  void run_batch(const std::vector<Task*>& batch, std::vector<Readback>* async);
  void fold_results(const std::vector<Task*>& batch, void* parts, void* tasks_dev, void* starts_dev, int maxn,
                    std::vector<Readback>* async);
  void distribute(Readback& rb);
  void* pinned_alloc(size_t bytes);
  void pinned_release(void* p);
  void* device_table(const std::string& bytes);  // immutable device copy of `bytes`, cached by content
  void* queue();                                  // the compute stream's chunk queue (zeroed once)
  void scalar_task(const Group& g);
  void ensure_input(Node* n);
  std::vector<Node*> d2_outputs(const Group& g, const std::vector<Node*>& targets);
  TileShape tile_for(const Group& g);
  void write_file(Node* w);

  // Plan cache: the fusion groups of an evaluation list and, per group and
  // set of materialised outputs, its generated kernel. A repeated evaluation
  // of the same sub-DAG (same nodes, same resident cut) reuses both (node ids
  // are never reused, so a matching key means the same live nodes).
  struct KernelEntry {
    KernelSpec spec;
    void* fn = nullptr;
    void* fin = nullptr;
    int occ = 0;  // resident work-groups per CU
  };
  struct Plan {
    std::vector<std::unique_ptr<Group>> groups;
    std::map<std::pair<const Group*, std::vector<Node*>>, KernelEntry> kernels;
  };
  std::string plan_key(const std::vector<Node*>& list, const std::vector<Node*>& targets) const;
  KernelEntry& kernel_for(Plan* plan, const Group& g, const std::vector<Node*>& kouts);
  // Kernels by shape (kernel_shape_key): a new group of a known shape (the same
  // script over another raster, the next request of a kind) takes the compiled
  // kernel and its parameter layout, its node references rebound by the local
  // numbering of kernel_order, instead of generating the source again.
  struct ShapeEntry {
    KernelEntry e;
    std::vector<int> param_node, d2_out, zout;  // local numbers of the spec's node references
  };
  std::unordered_map<std::string, ShapeEntry> shapes_;
  unsigned visit_stamp_ = 0;  // visit marks of cut_list and reach (Node::vstamp, Node::tstamp)
  std::vector<Node*> reach_, reach_stack_;
  const std::vector<Node*>& reach(const std::vector<Node*>& roots);  // nodes upwards of roots, each once
  std::unordered_map<std::string, std::shared_ptr<Plan>> plans_;
  std::multimap<size_t, void*> pinned_free_;
  std::unordered_map<void*, size_t> pinned_size_;
  std::string tune_key(const Group& g) const;
  std::map<std::string, TileShape> tuned_;
  std::map<std::string, std::string> kernel_group_;  // kernel name -> tune key
  std::map<std::string, std::string> kernel_src_;    // kernel name -> generated source
 public:
  std::map<std::string, int> kernel_occ_;            // kernel name -> resident work-groups per CU
 private:

  std::unordered_map<Node*, void*> resident_;
  std::unordered_map<Node*, int> pins_;  // rasters read by admitted requests: not dropped
  // Rasters moved out of HBM to pinned host memory when the store needed the
  // room (least recently used first); they come back when next read.
  struct Spilled {
    void* host = nullptr;
    bool generated = false;  // made by a generator: the host copy is its only source
  };
  std::unordered_map<Node*, Spilled> spilled_;
  unsigned long long use_clock_ = 0;               // one tick per evaluation / admitted request
  std::multiset<unsigned long long> live_ticks_;   // ticks of evaluations and requests in flight
  bool spill_one();        // moves the least recently used idle raster out of HBM
  void unspill(Node* n);   // brings a spilled raster back
  void forget_spilled(Node* n);
  unsigned long long tick_begin(const std::vector<Node*>& list);  // marks the nodes of a new evaluation
  void tick_end(unsigned long long t);
  struct ServeState;
  struct Request;
  struct CacheEntry {
    Node* node = nullptr;
    size_t bytes = 0;
    std::list<long long>::iterator pos;
  };
  std::list<long long> cache_lru_;  // most recent first
  std::unordered_map<long long, CacheEntry> cache_;
  size_t cache_budget_ = 0, cache_bytes_ = 0;
  long cache_hits_ = 0, cache_misses_ = 0, cache_evictions_ = 0;
  void cache_evict(long long key);
  std::unique_ptr<ServeState> srv_;
  void serve_retire();
  std::unordered_map<std::string, void*> tables_;
  size_t table_bytes_ = 0;
  void* queue_ = nullptr;
  // two pinned staging buffers through which files and host rasters stream
  // to and from HBM, each with the event of its last copy
  void* stage_buf_[2] = {nullptr, nullptr};
  void* stage_ev_[2] = {nullptr, nullptr};
  void* staging(int i);       // buffer i, once its last copy has completed
  void staged(int i);         // records the event of a copy just enqueued on buffer i
  std::recursive_mutex mtx_;
};

// Host evaluation of a D0 node from its (D0) prevs, with C semantics in the
// node types (ScalarTask, runtime/task/ScalarTask.cpp:69-193).
Variant eval_scalar_node(const Node* n);

}  // namespace mapx
