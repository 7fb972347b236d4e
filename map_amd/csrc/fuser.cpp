#include "fuser.hpp"
#include <algorithm>
#include <iterator>
#include <string>
#include <sstream>
#include <stdexcept>

namespace mapx {

bool pattern_is(int pat, int p) { return (pat & p) == p; }

namespace {
struct Rule { int a, b; bool r; };
// Pipe-fusion legality; the first matching row wins (runtime/Pattern.cpp:53-131).
const Rule kPipe[] = {
  {SPREAD,SPREAD,false},{SPREAD,RADIAL,false},{SPREAD,ZONAL,false},{SPREAD,FOCAL,false},
  {SPREAD,LOCAL,false},{SPREAD,FREE,true},{SPREAD,SPECIAL,false},{SPREAD,BARRIER,false},
  {RADIAL,SPREAD,false},{RADIAL,RADIAL,false},{RADIAL,ZONAL,false},{RADIAL,FOCAL,false},
  {RADIAL,LOCAL,true},{RADIAL,FREE,true},{RADIAL,SPECIAL,false},{RADIAL,BARRIER,false},
  {ZONAL,SPREAD,false},{ZONAL,RADIAL,false},{ZONAL,ZONAL,false},{ZONAL,FOCAL,false},
  {ZONAL,LOCAL,false},{ZONAL,FREE,true},{ZONAL,SPECIAL,true},{ZONAL,BARRIER,false},
  {FOCAL,SPREAD,false},{FOCAL,RADIAL,false},{FOCAL,ZONAL,false},{FOCAL,FOCAL,false},
  {FOCAL,LOCAL,true},{FOCAL,FREE,true},{FOCAL,SPECIAL,false},{FOCAL,BARRIER,false},
  {LOCAL,SPREAD,false},{LOCAL,RADIAL,true},{LOCAL,ZONAL,true},{LOCAL,FOCAL,false},
  {LOCAL,LOCAL,true},{LOCAL,FREE,true},{LOCAL,SPECIAL,false},{LOCAL,BARRIER,true},
  {SPECIAL,SPREAD,false},{SPECIAL,RADIAL,false},{SPECIAL,ZONAL,false},{SPECIAL,FOCAL,false},
  {SPECIAL,LOCAL,false},{SPECIAL,FREE,true},{SPECIAL,SPECIAL,false},{SPECIAL,BARRIER,false},
  {BARRIER,SPREAD,false},{BARRIER,RADIAL,false},{BARRIER,ZONAL,false},{BARRIER,FOCAL,false},
  {BARRIER,LOCAL,false},{BARRIER,FREE,true},{BARRIER,SPECIAL,false},{BARRIER,BARRIER,false},
};
// Flat-fusion legality (runtime/Pattern.cpp:133-182). A zero side matches any pattern.
const Rule kFlat[] = {
  {SPREAD,SPREAD,false},{SPREAD,RADIAL,false},{SPREAD,ZONAL,false},{SPREAD,FOCAL,false},
  {SPREAD,LOCAL,false},{SPREAD,FREE,true},
  {RADIAL,SPREAD,false},{RADIAL,RADIAL,false},{RADIAL,ZONAL,false},{RADIAL,FOCAL,false},
  {RADIAL,LOCAL,true},{RADIAL,FREE,true},
  {ZONAL,SPREAD,false},{ZONAL,RADIAL,false},{ZONAL,ZONAL,true},{ZONAL,FOCAL,false},
  {ZONAL,LOCAL,true},{ZONAL,FREE,true},
  {FOCAL,SPREAD,false},{FOCAL,RADIAL,false},{FOCAL,ZONAL,false},{FOCAL,FOCAL,true},
  {FOCAL,LOCAL,true},{FOCAL,FREE,true},
  {LOCAL,0,true},{0,LOCAL,true},{SPECIAL,0,true},{0,SPECIAL,true},
  {BARRIER,0,true},{0,BARRIER,true},{FREE,0,true},{0,FREE,true},
};
}  // namespace

// `top` is the producer's pattern on the edge, `bot` the consumer's (how the
// consuming nodes read the producer's output).
bool can_pipe_fuse(int top, int bot, bool focal_zonal) {
  // extension: a focal producer read cell by cell (not through a halo) by a
  // zonal consumer fuses into it
  if (focal_zonal && pattern_is(top, FOCAL) && pattern_is(bot, ZONAL) && !pattern_is(bot, FOCAL) &&
      !pattern_is(top, ZONAL) && !pattern_is(top, RADIAL) && !pattern_is(top, SPREAD) &&
      !pattern_is(bot, RADIAL) && !pattern_is(bot, SPREAD))
    return true;
  for (const Rule& r : kPipe)
    if (pattern_is(top, r.a) && pattern_is(bot, r.b)) return r.r;
  if (pattern_is(top, FREE) || pattern_is(bot, FREE)) return true;  // Pattern.cpp:122-123
  throw std::runtime_error("pipe-fusion: no rule for the pattern pair");
}

bool can_flat_fuse(int left, int right, bool focal_zonal) {
  if (focal_zonal && ((pattern_is(left, FOCAL) && pattern_is(right, ZONAL)) ||
                      (pattern_is(left, ZONAL) && pattern_is(right, FOCAL))))
    return !(pattern_is(left, RADIAL) || pattern_is(right, RADIAL) || pattern_is(left, SPREAD) ||
             pattern_is(right, SPREAD));
  for (const Rule& r : kFlat)
    if (pattern_is(left, r.a) && pattern_is(right, r.b)) return r.r;
  throw std::runtime_error("flat-fusion: no rule for the pattern pair");
}

// ---- Group -------------------------------------------------------------------

template <class T> static bool contains(const std::vector<T*>& v, const T* x) {
  return std::find(v.begin(), v.end(), x) != v.end();
}
template <class T> static void erase_value(std::vector<T*>& v, const T* x) {
  v.erase(std::remove(v.begin(), v.end(), x), v.end());
}

static int nd_rank(int nd) { return nd == D0 ? 0 : nd == D1 ? 1 : nd == D2 ? 2 : nd == D3 ? 3 : -1; }

Meta Group::meta() const {
  Meta best;
  int rank = -2;
  for (auto* lst : {&ins, &nodes})
    for (Node* n : *lst)
      if (nd_rank(n->meta.nd) > rank) { rank = nd_rank(n->meta.nd); best = n->meta; }
  return best;
}

std::string Group::signature(bool float_weights) const {
  std::ostringstream s;
  for (Node* n : ins) s << n->meta.nd << name_dt(n->meta.dt) << ";";
  s << "|";
  for (Node* n : nodes) s << n->signature(float_weights) << ";";
  s << "|";
  for (Node* n : outs) s << n->meta.nd << name_dt(n->meta.dt) << ";";
  return s.str();
}

void Group::addNode(Node* n) { if (!contains(nodes, n)) { nodes.push_back(n); pat |= n->pattern(); } }
void Group::removeNode(Node* n) { erase_value(nodes, n); }
void Group::addInput(Node* n) { if (!contains(ins, n)) ins.push_back(n); }
void Group::removeInput(Node* n) { erase_value(ins, n); }
void Group::addOutput(Node* n) { if (!contains(outs, n)) outs.push_back(n); }
void Group::removeOutput(Node* n) { erase_value(outs, n); }
void Group::addAuto(Node* n) {
  if (n->isInput()) { addInput(n); pat |= FREE; }
  else if (n->isOutput()) { addOutput(n); pat |= FREE; }
  else addNode(n);
}
void Group::removeAuto(Node* n) {
  if (n->isInput()) removeInput(n);
  else if (n->isOutput()) removeOutput(n);
  else removeNode(n);
}
void Group::addPrev(Group* g, int p) {
  auto it = std::find(prev.begin(), prev.end(), g);
  if (it != prev.end()) prev_pat[it - prev.begin()] |= p;
  else { prev.push_back(g); prev_pat.push_back(p); }
}
void Group::removePrev(Group* g) {
  auto it = std::find(prev.begin(), prev.end(), g);
  if (it == prev.end()) throw std::runtime_error("fuser: removePrev of a non-prev group");
  prev_pat.erase(prev_pat.begin() + (it - prev.begin()));
  prev.erase(it);
}
void Group::addNext(Group* g, int p) {
  auto it = std::find(next.begin(), next.end(), g);
  if (it != next.end()) next_pat[it - next.begin()] |= p;
  else { next.push_back(g); next_pat.push_back(p); }
}
void Group::removeNext(Group* g) {
  auto it = std::find(next.begin(), next.end(), g);
  if (it == next.end()) throw std::runtime_error("fuser: removeNext of a non-next group");
  next_pat.erase(next_pat.begin() + (it - next.begin()));
  next.erase(it);
}
int& Group::prevPattern(Group* g) {
  auto it = std::find(prev.begin(), prev.end(), g);
  if (it == prev.end()) throw std::runtime_error("fuser: prevPattern of a non-prev group");
  return prev_pat[it - prev.begin()];
}
int& Group::nextPattern(Group* g) {
  auto it = std::find(next.begin(), next.end(), g);
  if (it == next.end()) throw std::runtime_error("fuser: nextPattern of a non-next group");
  return next_pat[it - next.begin()];
}

// ---- Fuser -------------------------------------------------------------------

const std::vector<Node*>& Fuser::nexts(Node* n) const {
  static const std::vector<Node*> empty;
  auto it = next_.find(n);
  return it == next_.end() ? empty : it->second;
}

Group* Fuser::newGroup() {
  groups_.push_back(std::make_unique<Group>());
  return groups_.back().get();
}

void Fuser::removeGroup(Group* g) {
  groups_.erase(std::remove_if(groups_.begin(), groups_.end(),
                               [&](const std::unique_ptr<Group>& x) { return x.get() == g; }),
                groups_.end());
}

bool Fuser::reaches(const Group* from, const Group* to) const {
  // depth-first search over prev edges; groups are marked with a per-search stamp
  const unsigned st = ++stamp_;
  stack_.assign(from->prev.begin(), from->prev.end());
  while (!stack_.empty()) {
    const Group* g = stack_.back();
    stack_.pop_back();
    if (g == to) return true;
    if (g->stamp == st) continue;
    g->stamp = st;
    for (Group* p : g->prev)
      if (p->stamp != st) stack_.push_back(p);
  }
  return false;
}

// Groups over rasters of different geometry never share a kernel (one launch
// walks one block grid). The reference has no such check: its evaluations
// hold one raster geometry, while here independent scripts over rasters of
// different sizes may be evaluated together (and share D0 constants).
static bool same_geometry(const Group* a, const Group* b) {
  const Meta ma = a->meta(), mb = b->meta();
  if (ma.nd != D2 || mb.nd != D2) return true;
  return dim2_eq(ma.ds, mb.ds) && dim2_eq(ma.bs, mb.bs);
}

// Groups that run alone: radial and spreading scans (precompiled kernels,
// device/radial.hip and spread.hip, with no fused per-cell statements; the
// reference fuses LOCAL groups on either side, Pattern.cpp:71,95,149) and
// cell accesses (a D0 value read from a resident raster).
static bool solo(const Group* g) {
  if (pattern_is(g->pat, RADIAL) || pattern_is(g->pat, SPREAD)) return true;
  for (Node* n : g->nodes)
    if (n->kind == Kind::Access) return true;
  return false;
}

bool Fuser::canPipe(Group* top, Group* bot) {  // Fusioner.cpp:87-96
  if (top == bot) return false;
  if (solo(top) || solo(bot)) return false;
  if (!same_geometry(top, bot)) return false;
  for (Group* n : top->next)
    if (n != bot && isNext(bot, n)) return false;  // a path top -> n -> ... -> bot: fusing makes a cycle
  return can_pipe_fuse(bot->prevPattern(top), top->nextPattern(bot), opt_.focal_zonal) &&
         can_flat_fuse(top->pat, bot->pat, opt_.focal_zonal);
}

bool Fuser::canFlat(Group* left, Group* right) {  // Fusioner.cpp:98-106
  if (left == right) return false;
  if (solo(left) || solo(right)) return false;
  if (!same_geometry(left, right)) return false;
  if (isNext(right, left) || isNext(left, right)) return false;
  return can_flat_fuse(left->pat, right->pat, opt_.focal_zonal);
}

Group* Fuser::pipeFuse(Group* top, Group* bot) {  // Fusioner.cpp:108-183
  for (Node* n : bot->nodes) { top->addNode(n); erase_value(of_[n], bot); of_[n].push_back(top); }
  for (Node* n : bot->ins) { top->addInput(n); erase_value(of_[n], bot); of_[n].push_back(top); }
  for (Node* n : bot->outs) { top->addOutput(n); erase_value(of_[n], bot); of_[n].push_back(top); }
  for (Group* p : std::vector<Group*>(bot->prev)) {
    if (p == top) continue;
    p->addNext(top, p->nextPattern(bot));
    top->addPrev(p, bot->prevPattern(p));
    p->removeNext(bot);
  }
  for (Group* n : std::vector<Group*>(bot->next)) {
    if (n == top) throw std::runtime_error("fuser: cycle in pipe fusion");
    n->addPrev(top, n->prevPattern(bot) | bot->prevPattern(top));
    top->addNext(n, bot->nextPattern(n));
    n->removePrev(bot);
  }
  for (Group* p : top->prev) p->addNext(top, top->nextPattern(bot));
  top->removeNext(bot);
  top->pat |= bot->pat;
  removeGroup(bot);
  return top;
}

Group* Fuser::flatFuse(Group* left, Group* right) {  // Fusioner.cpp:185-245
  for (Node* n : right->nodes) { left->addNode(n); erase_value(of_[n], right); of_[n].push_back(left); }
  for (Node* n : right->ins) { left->addInput(n); erase_value(of_[n], right); of_[n].push_back(left); }
  for (Node* n : right->outs) { left->addOutput(n); erase_value(of_[n], right); of_[n].push_back(left); }
  for (Group* p : std::vector<Group*>(right->prev)) {
    p->addNext(left, p->nextPattern(right));
    left->addPrev(p, right->prevPattern(p));
    p->removeNext(right);
  }
  for (Group* n : std::vector<Group*>(right->next)) {
    n->addPrev(left, n->prevPattern(right));
    left->addNext(n, right->nextPattern(n));
    n->removePrev(right);
  }
  left->pat |= right->pat;
  removeGroup(right);
  return left;
}

void Fuser::process(Node* node) {  // Fusioner.cpp:247-257
  Group* g = newGroup();
  g->addAuto(node);
  of_[node].push_back(g);
  for (Node* p : node->prev) {
    if (!in_eval_.count(p)) continue;
    Group* pg = of_[p].front();
    pg->addNext(g, g->pat);
    g->addPrev(pg, pg->pat);
  }
}

static bool free_or_local(const Group* g) { return pattern_is(g->pat, FREE) || pattern_is(g->pat, LOCAL); }

void Fuser::pipeGently(Node* node) {  // Fusioner.cpp:259-278
  if (!opt_.code_fusion) return;
  Group* g = of_[node].front();
  size_t i = 0;
  while (i < g->prev.size()) {
    Group* pg = g->prev[i++];
    bool fuse_free = free_or_local(g) && free_or_local(pg);
    bool fuse_dnd0 = !(g->numdim() != D0 && pg->numdim() == D0 && pg->pat != FREE);
    if (fuse_free && fuse_dnd0 && pg->next.size() == 1 && canPipe(pg, g)) {
      g = pipeFuse(pg, g);
      i = 0;
    }
  }
}

void Fuser::flatGently(Node* node) {  // Fusioner.cpp:280-309
  if (!opt_.code_fusion) return;
  Group* ng = of_[node].front();
  if (ng->next.size() < 2) return;
  const auto& nx = nexts(node);
  // candidates of one left group: the later next-nodes whose groups have its
  // raster geometry (same_geometry rejects every other pair), in order
  std::map<std::string, std::vector<size_t>> by_geom;
  std::vector<std::string> key(nx.size());
  for (size_t k = 0; k < nx.size(); k++) {
    const Meta m = of_[nx[k]].front()->meta();
    if (m.nd == D2) {
      int v[4] = {m.ds.arr[0], m.ds.arr[1], m.bs.arr[0], m.bs.arr[1]};
      key[k].assign((const char*)v, sizeof v);
    }
    by_geom[key[k]].push_back(k);
  }
  size_t i = 0;
  while (i < nx.size()) {
    const size_t li = i;
    Group* left = of_[nx[i++]].front();
    if (!free_or_local(left) || left == ng) continue;
    // a D0 left group may pair with any geometry: scan all later nodes
    const bool any = key[li].empty();
    std::vector<size_t> cand;
    if (any) for (size_t j = i; j < nx.size(); j++) cand.push_back(j);
    else {
      const auto& same = by_geom[key[li]];
      const auto& d0 = by_geom[std::string()];
      std::merge(std::upper_bound(same.begin(), same.end(), li), same.end(), std::upper_bound(d0.begin(), d0.end(), li),
                 d0.end(), std::back_inserter(cand));
    }
    for (size_t j : cand) {
      Group* right = of_[nx[j]].front();
      if (!free_or_local(right) || right == ng || right == left) continue;
      if (canFlat(left, right)) left = flatFuse(left, right);
    }
  }
}

void Fuser::processBU(Group* g) {  // Fusioner.cpp:340-392
  if (!opt_.code_fusion) return;
  if (!visited_.insert(g).second) return;
  size_t i = 0;
  while (i < g->prev.size()) {
    Group* bot = g;
    Group* top = g->prev[i++];
    bool d0dn = !(top->pat != FREE && top->numdim() == D0 && bot->numdim() != D0);
    if (d0dn && canPipe(top, bot)) {
      g = pipeFuse(top, bot);
      i = 0;
    }
  }
  size_t s = g->prev.size();
  i = 0;
  while (i < s) {
    processBU(g->prev[i++]);
    if (s != g->prev.size()) { s = g->prev.size(); i = 0; }
  }
}

std::vector<Node*> Fuser::allNodes(Group* g) const {
  std::vector<Node*> v(g->ins);
  v.insert(v.end(), g->nodes.begin(), g->nodes.end());
  v.insert(v.end(), g->outs.begin(), g->outs.end());
  return v;
}

void Fuser::forwarding() {  // Fusioner.cpp:394-450, predicate: FREE nodes
  std::map<std::pair<Group*, Group*>, std::vector<Node*>> forward;
  // after its first visit a FREE node is in the group of every next-node
  // (groups do not change here), so later visits would add nothing
  std::set<Node*> visited;
  for (size_t gi = 0; gi < groups_.size(); gi++) {
    Group* g = groups_[gi].get();
    for (Node* node : allNodes(g)) {
      if (node->pattern() != FREE) continue;
      if (!visited.insert(node).second) continue;
      bool forwarded = false;
      for (Node* nn : nexts(node)) {
        for (Group* ng : std::vector<Group*>(of_[nn])) {
          auto all = allNodes(ng);
          if (!contains(all, node)) {
            ng->addAuto(node);
            of_[node].push_back(ng);
            forwarded = true;
            forward[{g, ng}].push_back(node);
          }
        }
      }
      if (forwarded) {
        bool none_inside = true;
        for (Node* nn : nexts(node)) if (contains(g->nodes, nn)) none_inside = false;
        if (none_inside) { g->removeAuto(node); erase_value(of_[node], g); }
      }
    }
  }
  for (auto& f : forward) {
    Group* g = f.first.first;
    Group* ng = f.first.second;
    bool linked = false;
    for (Node* n : g->nodes) {
      if (contains(f.second, n) || n->pattern() == FREE) continue;
      for (Node* nn : nexts(n)) if (contains(ng->nodes, nn)) linked = true;
    }
    if (!linked) {
      if (contains(g->next, ng)) g->removeNext(ng);
      if (contains(ng->prev, g)) ng->removePrev(g);
    }
  }
  groups_.erase(std::remove_if(groups_.begin(), groups_.end(),
                               [&](const std::unique_ptr<Group>& x) {
                                 return x->ins.empty() && x->nodes.empty() && x->outs.empty();
                               }),
                groups_.end());
}

void Fuser::linking() {  // Fusioner.cpp:452-509
  // a node computed in group g and read in group ng (which does not compute
  // it) becomes an output of g and an input of ng; visited per node, so a
  // node computed in many groups (a forwarded constant) costs its edges once
  std::unordered_map<Node*, std::vector<Group*>> in_nodes;
  for (auto& up : groups_)
    for (Node* n : up->nodes) in_nodes[n].push_back(up.get());
  for (auto& e : in_nodes) {
    Node* n = e.first;
    const std::vector<Group*>& gs = e.second;
    for (Node* nn : nexts(n))
      for (Group* ng : of_[nn])
        if (!contains(gs, ng))
          for (Group* g : gs) { g->addOutput(n); ng->addInput(n); }
  }
  // D0 FREE nodes become inputs: better reuse of compiled kernels (Fusioner.cpp:494-508).
  for (auto& up : groups_) {
    Group* g = up.get();
    size_t i = 0;
    while (i < g->nodes.size()) {
      Node* n = g->nodes[i];
      if (n->meta.nd == D0 && n->pattern() == FREE) { g->removeNode(n); g->addInput(n); }
      else i++;
    }
  }
}

static bool node_less(const Node* a, const Node* b) {
  if (a->id != b->id) return a->id < b->id;
  return (int)a->isOutput() < (int)b->isOutput();
}

void Fuser::sorting() {  // Fusioner.cpp:511-543
  for (auto& g : groups_) {
    std::sort(g->nodes.begin(), g->nodes.end(), node_less);
    std::sort(g->ins.begin(), g->ins.end(), node_less);
    std::sort(g->outs.begin(), g->outs.end(), node_less);
  }
  // Topological order: a group precedes every group that depends on it; ties
  // broken by the id of the last node. A stable topological sort replaces the
  // reference's std::sort with a partial-order comparator.
  std::unordered_map<Group*, int> indeg, last;
  for (auto& g : groups_) {
    indeg[g.get()] = (int)g->prev.size();
    int m = -1;
    for (Node* n : g->nodes) m = std::max(m, n->id);
    for (Node* n : g->outs) m = std::max(m, n->id);
    last[g.get()] = m;
  }
  // ready groups by (last node id, position): the earliest-finishing first
  std::set<std::pair<int, size_t>> ready;
  std::unordered_map<Group*, size_t> pos;
  for (size_t i = 0; i < groups_.size(); i++) pos[groups_[i].get()] = i;
  for (auto& g : groups_) if (indeg[g.get()] == 0) ready.insert({last[g.get()], pos[g.get()]});
  std::vector<std::unique_ptr<Group>> sorted;
  while (!ready.empty()) {
    const size_t i = ready.begin()->second;
    ready.erase(ready.begin());
    Group* g = groups_[i].get();
    sorted.push_back(std::move(groups_[i]));
    for (Group* n : g->next) if (--indeg[n] == 0) ready.insert({last[n], pos[n]});
  }
  if (sorted.size() != groups_.size()) throw std::runtime_error("fuser: group graph has a cycle");
  groups_.swap(sorted);
  int id = 0;
  for (auto& g : groups_) g->id = id++;
}

std::vector<std::unique_ptr<Group>> Fuser::fuse(const std::vector<Node*>& list) {
  groups_.clear(); of_.clear(); next_.clear(); visited_.clear();
  in_eval_ = std::set<Node*>(list.begin(), list.end());
  for (Node* n : list) {
    std::vector<Node*> v;
    for (Node* m : n->next) if (in_eval_.count(m)) v.push_back(m);
    std::sort(v.begin(), v.end(), node_less);
    next_[n] = v;
  }
  for (Node* n : list) { process(n); pipeGently(n); }
  for (Node* n : list) flatGently(n);
  for (auto it = list.rbegin(); it != list.rend(); ++it) {
    if (of_[*it].size() != 1) throw std::runtime_error("fuser: node in more than one group before forwarding");
    processBU(of_[*it].front());
  }
  forwarding();
  linking();
  sorting();
  return std::move(groups_);
}

}  // namespace mapx
