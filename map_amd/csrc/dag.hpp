// Lazy DAG of map-algebra operations with online global value numbering.
//
// Node kinds and their typing rules follow runtime/dag/*.cpp of the reference
// (factories cited per kind in dag.cpp). A node is immutable once added; the
// value-numbering table returns an existing equal node instead of a new one
// (runtime/visitor/SimplifierOnline.tpl:13-38).
#pragma once
#include "types.hpp"
#include <functional>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace mapx {

enum class Kind : int {
  Constant, Index, Cast, Unary, Binary, Conditional,
  Neighbor, Convolution, FocalFunc,
  ZonalReduc, ZonalStats,
  Read, Write, Scalar,
  RadialScan, FocalPercent, Diversity,
  SpreadNeighbor, SpreadScan, Stats, Barrier, Access, LhsAccess, Rand
};
const char* kind_name(Kind k);

struct Meta {
  int dt = NONE_DATATYPE;
  int nd = NONE_NUMDIM;
  Dim2 ds = dim2_none();
  int mo = NONE_MEMORDER;
  Dim2 bs = dim2_none();
  int nb(int i) const { return (ds.arr[i] + bs.arr[i] - 1) / bs.arr[i]; }  // MetaData.hpp:59
};

// Focal mask: ds = {ncols, nrows}, values row-major (python/map.py:542-550),
// dtype = dtype of the first value (util/Mask.cpp:18).
struct Mask {
  Dim2 ds = dim2_none();
  std::vector<Variant> v;
  int dtype() const { return v.empty() ? NONE_DATATYPE : v[0].dtype; }
  int halo(int i) const { return ds.arr[i] / 2; }
  const Variant& at(int x, int y) const { return v[(size_t)y * ds.arr[0] + x]; }
};

struct HostRaster;  // in-memory raster handed in through the extension ABI

struct Node {
  int id = -1;
  Kind kind;
  Meta meta;
  std::vector<Node*> prev, next;
  // nodes evaluated together with this one and kept alive by it, without a data
  // dependency (the MAX / MIN reductions of a Stats node)
  std::vector<Node*> aux;
  int ref = 0;  // references held by the front end (ma_increaseRef)
  std::string gkey;  // value-numbering key (set when the node enters the DAG)

  // kind parameters
  int op = 0;            // Unary/Binary/Reduction enum, Index dim, Cast target
  Variant cnst{};        // Constant
  Mask mask;             // Convolution / FocalFunc
  Dim2 coord = dim2_none();  // Neighbor offset; RadialScan start cell; Access / LhsAccess cell
  std::string path;      // Read / Write
  int nzones = 0;        // ZonalStats
  std::shared_ptr<HostRaster> host;  // Read of an in-memory raster

  // results
  Variant value{};       // D0 result after evaluation
  bool evaluated = false;
  std::vector<unsigned char> table;  // ZonalStats result: nzones records of 24 bytes
  // Set during one evaluation for a node whose value is already resident in
  // HBM (D2) or known (D0): it enters the evaluation as an input.
  bool cut = false;
  unsigned vstamp = 0, tstamp = 0;  // per-listing visit / target marks of the runtime
  unsigned long long use = 0;        // the runtime's tick of the last evaluation that listed the node
  int pattern() const;
  bool isInput() const { return kind == Kind::Read || cut; }
  bool isOutput() const { return kind == Kind::Write || kind == Kind::Scalar; }
  int halo(int i) const;
  // kind, type and parameters; float_weights = false leaves out floating mask
  // weights (kernels take them as arguments)
  std::string signature(bool float_weights = true) const;
};

// In-memory raster registered by ma_input: block-tiled (ROW+BLK) bytes, the
// same layout as the payload of a .bin file.
struct HostRaster {
  Meta meta;
  std::vector<unsigned char> bytes;  // host copy (may be empty when dev is set)
  void* dev = nullptr;               // device raster made by a generator (runtime-owned)
  bool dev_owned = false;
  size_t nbytes = 0;
  uint64_t uid = 0;
};

class Dag {
 public:
  // Factories with the reference's type/shape inference. Each returns the
  // value-numbered node (an existing equal node when one exists).
  Node* constant(Variant v, Dim2 ds, int dt, int mo, Dim2 bs);
  Node* index(Node* p, int dim);
  Node* rand(Node* seed, int dt, int mo);
  Node* cast(Node* p, int dt);
  Node* unary(Node* p, int op);
  Node* binary(Node* l, Node* r, int op);
  Node* conditional(Node* c, Node* l, Node* r);
  Node* neighbor(Node* p, Dim2 coord);
  Node* convolution(Node* p, const Mask& m);
  Node* focalFunc(Node* p, const Mask& m, int red);
  Node* zonalReduc(Node* p, int red);
  Node* zonalStats(Node* data, Node* zone, int nzones);
  Node* radialScan(Node* p, int red, Dim2 start);
  Node* focalPercent(Node* p, const Mask& m, int type);
  Node* diversity(const std::vector<Node*>& ps, int type);
  Node* spreadNeighbor(Node* p, Node* dir, int red);
  Node* spreadScan(Node* p, Node* dir, int red);
  Node* stats(Node* p);
  Node* barrier(Node* p);
  Node* access(Node* p, Dim2 cell);
  Node* lhsAccess(Node* lhs, Node* rhs, Dim2 cell);
  Node* read(const std::string& path);
  Node* input(std::shared_ptr<HostRaster> r);
  Node* write(Node* p, const std::string& path);
  Node* scalar(Node* p);

  // Drops nodes that nobody references and that no referenced node depends
  // on (Runtime::unlinkIsolated, runtime/Runtime.cpp:375-393).
  // `on_delete` is called for every node before it is destroyed.
  void collect(const std::function<void(Node*)>& on_delete = nullptr);
  // collect() once the node list has doubled since the last collection (and
  // holds at least 4096 nodes): callers that run every evaluation or serving
  // step pay O(1) per created node instead of a full scan per call. Rasters
  // of unreferenced nodes are released separately and at once.
  void collect_grown(const std::function<void(Node*)>& on_delete = nullptr);
  size_t size() const { return nodes_.size(); }
  const std::vector<std::unique_ptr<Node>>& nodes() const { return nodes_; }

 private:
  Node* add(std::unique_ptr<Node> n);
  std::string key(const Node* n) const;
  std::vector<std::unique_ptr<Node>> nodes_;
  std::unordered_map<std::string, Node*> gvn_;
  size_t live_after_collect_ = 0;
  int id_count_ = 0;
};

// Sub-DAG reachable (upwards) from the given nodes, sorted by id
// (runtime/visitor/ListerBU + runtime/Runtime.cpp:426-440).
std::vector<Node*> list_upwards(const std::vector<Node*>& roots);

}  // namespace mapx
