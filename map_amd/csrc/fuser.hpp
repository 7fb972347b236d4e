// Loop-fusion compiler: groups the nodes of one evaluation into fused groups.
//
// The stages, their order and the legality tables restate
// runtime/visitor/Fusioner.cpp:34-72 and runtime/Pattern.cpp:53-258, because
// the grouping decides which kernels exist and where the block-relative focal
// border rule applies: a different grouping gives different results at block
// seams. Two opt-in extensions exist: FOCAL->ZONAL pipe fusion (the fused
// hillshade + zonal form) and FOCAL->FOCAL is left as in the reference (off).
#pragma once
#include "dag.hpp"
#include <map>
#include <memory>
#include <set>
#include <unordered_map>
#include <vector>

namespace mapx {

bool pattern_is(int pat, int p);  // (pat & p) == p, runtime/Pattern.cpp:46-48
bool can_pipe_fuse(int top, int bot, bool focal_zonal);
bool can_flat_fuse(int left, int right, bool focal_zonal);

struct Group {
  int id = -1;
  std::vector<Node*> nodes, ins, outs;
  int pat = NONE_PAT;
  std::vector<Group*> prev, next;
  std::vector<int> prev_pat, next_pat;
  mutable unsigned stamp = 0;  // visit mark of reachability searches

  Meta meta() const;  // metadata of the highest-dimensional input or node (Group.cpp:143-156)
  int numdim() const { return meta().nd; }
  std::string signature(bool float_weights = true) const;  // Group.cpp:262-272

  void addNode(Node* n);
  void removeNode(Node* n);
  void addInput(Node* n);
  void removeInput(Node* n);
  void addOutput(Node* n);
  void removeOutput(Node* n);
  void addAuto(Node* n);     // Group.cpp:116-127
  void removeAuto(Node* n);
  void addPrev(Group* g, int pat);
  void removePrev(Group* g);
  void addNext(Group* g, int pat);
  void removeNext(Group* g);
  int& prevPattern(Group* g);
  int& nextPattern(Group* g);
};

struct FuseOptions {
  bool code_fusion = true;  // Config::code_fusion
  bool focal_zonal = false; // extension: fuse FOCAL -> ZONAL into one group
};

class Fuser {
 public:
  explicit Fuser(FuseOptions o) : opt_(o) {}
  // `list` is the evaluation's sub-DAG sorted by id; edges to nodes outside it
  // are ignored (the reference clones the sub-DAG, runtime/Runtime.cpp:448-452).
  std::vector<std::unique_ptr<Group>> fuse(const std::vector<Node*>& list);

  const std::vector<Node*>& nexts(Node* n) const;  // next-nodes within the evaluation

 private:
  Group* newGroup();
  void removeGroup(Group* g);
  bool reaches(const Group* from, const Group* to) const;  // `to` is a transitive prev of `from`
  bool isPrev(const Group* g, const Group* of) const { return reaches(of, g); }
  bool isNext(const Group* g, const Group* of) const { return reaches(g, of); }
  bool canPipe(Group* top, Group* bot);
  bool canFlat(Group* left, Group* right);
  Group* pipeFuse(Group* top, Group* bot);
  Group* flatFuse(Group* left, Group* right);
  void process(Node* n);
  void pipeGently(Node* n);
  void flatGently(Node* n);
  void processBU(Group* g);
  void forwarding();
  void linking();
  void sorting();
  std::vector<Node*> allNodes(Group* g) const;

  FuseOptions opt_;
  std::vector<std::unique_ptr<Group>> groups_;
  std::unordered_map<Node*, std::vector<Group*>> of_;
  std::unordered_map<Node*, std::vector<Node*>> next_;
  std::set<Group*> visited_;
  std::set<Node*> in_eval_;
  mutable unsigned stamp_ = 0;
  mutable std::vector<const Group*> stack_;
};

}  // namespace mapx
