// Host-side solver-order authority: the accurate and speculative island
// graphs of the reference's native scene, kept incrementally from each
// step's broadphase and touch changes, and the solver order they imply
// (the active accurate islands' contact-manager arrays, batched into
// solver tasks of at least 128 bodies and one constraint).
//
// The reference keeps these graphs on the host in its GPU mode too
// (scene.cpp:5119-5952 update_island_connectivity, 4759-5118 lost-edge
// routes and splits, 6148-6262 wakes, 6465-6536 sleeps, 6746-6884 active
// lists, 7174-7260 task batching). The order is a function of the whole
// touch history: merges keep the larger island and concatenate contact
// arrays, lost edges split islands along hop-count routes, removals swap
// the last array entry into the hole, island handles come back from a
// free stack. Every rule that decides an order is restated here; nothing
// else of the scene (static touch counters, kinematic nodes, joints, actor
// retirement, rollback journals) is.
//
// Scope: dynamic bodies and static bodies, one shape per body, bodies
// admitted before the first step and never removed.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "fyzz_hip.h"

namespace fyzz_hip::islands {

using u32 = std::uint32_t;
using u64 = std::uint64_t;
constexpr u32 kNone = 0xFFFFFFFFu;
constexpr std::size_t kTaskBodies = 128;  // solver batch size (scene.cpp:7174)

inline u64 key(u32 a, u32 b) { return (u64(a) << 32) | b; }

struct Link {
  u32 prev = kNone, next = kNone;  // instance = edge * 2 + side
  bool linked = false;
};

// One contact manager: a broadphase pair of bodies (a < b). Its slot is
// allocated when the pair is found and returned to the free stack after
// the step that loses it (scene.cpp:2748-2790, 3024-3041).
struct Edge {
  u32 a = kNone, b = kNone;
  Link link[2][2];      // [graph][side]: side 0 sits in a's list, 1 in b's
  u32 slot = kNone;     // position in its accurate island's contact array
  bool live = false, retiring = false, connected = false;
};

// One entry of an island's contact array: the manager and its bodies (so
// writing the solver order reads the array alone).
struct Contact {
  u32 id, a, b;
};

struct Island {
  u32 root = kNone, tail = kNone, count = 0;
  u32 active = kNone;            // index in the active island list
  bool live = false;
  std::vector<Contact> contacts; // accurate graph only
};

class Authority;

// Broadphase pair -> manager slot: open addressing with linear probing and
// backward-shift erase, kept at most half full. A key is (a << 32) | b with
// a < b, so the all-ones key never names a pair and marks an empty cell.
class PairMap {
 public:
  explicit PairMap(std::size_t expected) { rehash(expected * 2); }
  u32 find(u64 k) const {
    for (std::size_t i = home(k);; i = (i + 1) & mask_) {
      if (cells_[i].key == k) return cells_[i].value;
      if (cells_[i].key == kEmpty) return kNone;
    }
  }
  void insert(u64 k, u32 v) {
    if (2 * (size_ + 1) > cells_.size()) rehash(cells_.size() * 2);
    std::size_t i = home(k);
    while (cells_[i].key != kEmpty && cells_[i].key != k) i = (i + 1) & mask_;
    if (cells_[i].key == kEmpty) ++size_;
    cells_[i] = Cell{k, v};
  }
  void erase(u64 k) {
    std::size_t i = home(k);
    while (cells_[i].key != k) {
      if (cells_[i].key == kEmpty) return;
      i = (i + 1) & mask_;
    }
    // pull later entries of the probe run back over the hole
    for (std::size_t j = (i + 1) & mask_;; j = (j + 1) & mask_) {
      if (cells_[j].key == kEmpty) break;
      const std::size_t h = home(cells_[j].key);
      if (((j - h) & mask_) >= ((j - i) & mask_)) {
        cells_[i] = cells_[j];
        i = j;
      }
    }
    cells_[i].key = kEmpty;
    --size_;
  }

 private:
  static constexpr u64 kEmpty = ~u64(0);
  struct Cell {
    u64 key = kEmpty;
    u32 value = kNone;
  };
  std::size_t home(u64 k) const { return std::size_t((k * 0x9E3779B97F4A7C15ull) >> 20) & mask_; }
  void rehash(std::size_t want) {
    std::size_t cap = 1024;
    while (cap < want) cap *= 2;
    std::vector<Cell> old(cap);
    old.swap(cells_);
    mask_ = cap - 1;
    size_ = 0;
    for (const Cell& c : old)
      if (c.key != kEmpty) insert(c.key, c.value);
  }
  std::vector<Cell> cells_;
  std::size_t mask_ = 0, size_ = 0;
};

// One island graph (0: speculative, over broadphase pairs; 1: accurate,
// over touching managers).
struct Graph {
  int kind = 0;
  std::vector<Island> islands;
  std::vector<u32> free_handles;
  std::vector<u32> node_island, node_next, node_prev;
  std::vector<u32> head;                     // node -> first instance
  std::vector<u32> hops, fast, fast_edge;
  std::vector<uint8_t> hint;                 // route hint is current
  std::vector<u32> active_nodes, active_index;
  std::size_t initial_active = 0;
  std::vector<u32> active_islands;
  std::vector<u32> activating;
  std::vector<u32> dirty;
  std::vector<uint8_t> dirty_mark;
  // route search scratch: one record per node, so a search reads one
  // line per node it reaches; heap entries carry their node's hop estimate
  // (fixed while a search runs) above the node
  struct Mark {
    u32 seen = 0, pred = kNone, pred_edge = kNone, confirmed = 0;
  };
  std::vector<Mark> mark;
  u32 confirm_stamp = 0, visit_stamp = 0;
  std::vector<u64> heap;
  std::vector<u32> members, split_edges;
  std::vector<u32> island_seen, expand_seen;
  u32 island_stamp = 0, expand_stamp = 0;

  void resize(std::size_t n) {
    node_island.assign(n, kNone);
    node_next.assign(n, kNone);
    node_prev.assign(n, kNone);
    head.assign(n, kNone);
    hops.assign(n, 0);
    fast.assign(n, kNone);
    fast_edge.assign(n, kNone);
    hint.assign(n, 0);
    active_index.assign(n, kNone);
    dirty_mark.assign(n, 0);
    mark.assign(n, Mark{});
  }
};

class Authority {
 public:
  Authority(u32 n, const uint8_t* dynamic) : n_(n), dyn_(dynamic, dynamic + n), by_pair_(4ull * n) {
    sleeping_.assign(n, 0);
    ready_.assign(n, 0);
    for (int g = 0; g < 2; ++g) {
      graph_[g].kind = g;
      graph_[g].resize(n);
    }
  }

  void publish(const fyzz_island_events& ev, fyzz_island_order& out);
  const std::vector<uint8_t>& sleeping() const { return sleeping_; }

 private:
  bool dynamic(u32 node) const { return node < n_ && dyn_[node]; }
  u32 edge_of(u32 a, u32 b) const {
    return by_pair_.find(key(a, b));
  }
  u32 owner_node(const Edge& e) const { return dynamic(e.a) ? e.a : e.b; }

  Link& instance(int g, u32 inst) { return edges_[inst / 2].link[g][inst & 1]; }
  void link(int g, u32 id);
  void unlink(int g, u32 id);
  void append_contact(u32 id, u32 owner);
  void remove_contact(u32 id);

  void admit(Graph& G);
  bool merge(Graph& G, u32 a, u32 b, u32 edge);
  void rebuild_chains(Graph& G);
  bool repair(Graph& G);
  void split(Graph& G, u32 seed);
  void heap_push(Graph& G, u32 node);
  u32 heap_pop(Graph& G);
  bool live_hop(const Graph& G, u32 from, u32 to, u32 id) const;
  void synchronize(Graph& G);
  void activate(const fyzz_island_events& ev);
  void deactivate(Graph& G, std::vector<u32>& went_to_sleep);

  u32 n_;
  std::vector<uint8_t> dyn_, sleeping_, ready_;
  std::vector<Edge> edges_;
  std::vector<u32> free_edges_, retired_edges_;
  PairMap by_pair_;  // broadphase pairs: a few per body
  Graph graph_[2];
  bool admitted_ = false;
};

// Push-front into both endpoint lists of graph g (scene.cpp:2823-2851).
void Authority::link(int g, u32 id) {
  Graph& G = graph_[g];
  Edge& e = edges_[id];
  for (int side = 0; side < 2; ++side) {
    const u32 node = side == 0 ? e.a : e.b;
    if (!dynamic(node)) continue;  // a static endpoint has no island node
    Link& l = e.link[g][side];
    if (l.linked) continue;
    const u32 inst = id * 2 + side;
    const u32 first = G.head[node];
    l.prev = kNone;
    l.next = first;
    l.linked = true;
    if (first != kNone) instance(g, first).prev = inst;
    G.head[node] = inst;
  }
}

// Remove from both endpoint lists; a dynamic/dynamic edge marks both
// endpoints dirty for the lost-edge pass and drops a route hint that
// used it (scene.cpp:2937-3023).
void Authority::unlink(int g, u32 id) {
  Graph& G = graph_[g];
  if (g == 1 && edges_[id].connected) remove_contact(id);
  Edge& e = edges_[id];
  for (int side = 0; side < 2; ++side) {
    const u32 node = side == 0 ? e.a : e.b;
    Link& l = e.link[g][side];
    if (!l.linked) continue;
    const u32 other = side == 0 ? e.b : e.a;
    if (dynamic(node) && dynamic(other)) {
      if (!G.dirty_mark[node]) {
        G.dirty.push_back(node);
        G.dirty_mark[node] = 1;
      }
      if (G.fast[node] == other) G.fast[node] = kNone;
    }
    if (l.prev == kNone) G.head[node] = l.next;
    else instance(g, l.prev).next = l.next;
    if (l.next != kNone) instance(g, l.next).prev = l.prev;
    l = Link{};
  }
}

// An island's contact managers form an array: appends go to the end,
// a removal moves the last entry into the hole (scene.cpp:2855-2934).
void Authority::append_contact(u32 id, u32 owner) {
  Island& isl = graph_[1].islands[owner];
  Edge& e = edges_[id];
  if (!isl.live || !e.connected || e.slot != kNone)
    throw std::logic_error("island contact insertion out of order");
  e.slot = static_cast<u32>(isl.contacts.size());
  isl.contacts.push_back({id, e.a, e.b});
}

void Authority::remove_contact(u32 id) {
  Edge& e = edges_[id];
  if (e.slot == kNone) return;
  const u32 owner = graph_[1].node_island[owner_node(e)];
  if (owner == kNone) return;
  auto& c = graph_[1].islands[owner].contacts;
  if (e.slot >= c.size() || c[e.slot].id != id) throw std::logic_error("island contact array corrupt");
  const Contact last = c.back();
  c[e.slot] = last;
  edges_[last.id].slot = e.slot;
  c.pop_back();
  e.slot = kNone;
}

// Every dynamic body not yet in the graph becomes its own island; awake
// ones queue for activation (scene.cpp:5352-5407).
void Authority::admit(Graph& G) {
  for (u32 id = 0; id < n_; ++id) {
    if (!dynamic(id) || G.node_island[id] != kNone) continue;
    u32 h;
    if (!G.free_handles.empty()) {
      h = G.free_handles.back();
      G.free_handles.pop_back();
    } else {
      h = static_cast<u32>(G.islands.size());
      G.islands.emplace_back();
    }
    Island& isl = G.islands[h];
    isl = Island{};
    isl.root = isl.tail = id;
    isl.count = 1;
    isl.live = true;
    G.node_island[id] = h;
    G.node_next[id] = G.node_prev[id] = kNone;
    if (!sleeping_[id]) G.activating.push_back(id);
  }
}

// A new edge joins two islands: the larger keeps its handle (ties keep
// the second endpoint's), the smaller's nodes follow its chain and its
// contact array follows the winner's; hop estimates of the absorbed nodes
// grow by the joining path. Inside one island only a shorter hop route is
// recorded (scene.cpp:5233-5346).
bool Authority::merge(Graph& G, u32 pa, u32 pb, u32 edge) {
  if (!dynamic(pa) || !dynamic(pb)) return false;
  const u32 a = G.node_island[pa], b = G.node_island[pb];
  if (a == b) {
    const u32 ha = G.hops[pa], hb = G.hops[pb];
    if (ha + 1 < hb) {
      G.hops[pb] = ha + 1;
      G.fast[pb] = pa;
      G.fast_edge[pb] = edge;
      G.hint[pb] = 1;
    } else if (hb + 1 < ha) {
      G.hops[pa] = hb + 1;
      G.fast[pa] = pb;
      G.fast_edge[pa] = edge;
      G.hint[pa] = 1;
    }
    return false;
  }
  const bool keep_a = G.islands[a].count > G.islands[b].count;
  const u32 kept = keep_a ? a : b, gone = keep_a ? b : a;
  const u32 wnode = keep_a ? pa : pb, lnode = keep_a ? pb : pa;
  Island& w = G.islands[kept];
  Island& l = G.islands[gone];
  const u32 extra = G.hops[wnode] + G.hops[lnode] + 1;
  for (u32 v = l.root; v != kNone; v = G.node_next[v]) {
    G.hops[v] += extra;
    G.node_island[v] = kept;
  }
  G.hops[lnode] = G.hops[wnode] + 1;
  G.fast[lnode] = wnode;
  G.fast_edge[lnode] = edge;
  G.hint[lnode] = 1;
  G.node_prev[l.root] = w.tail;
  G.node_next[w.tail] = l.root;
  w.tail = l.tail;
  w.count += l.count;
  if (G.kind == 1) {
    for (const Contact& c : l.contacts) {
      edges_[c.id].slot = static_cast<u32>(w.contacts.size());
      w.contacts.push_back(c);
    }
  }
  if (l.active != kNone) {
    const u32 idx = l.active, moved = G.active_islands.back();
    G.active_islands[idx] = moved;
    G.islands[moved].active = idx;
    G.active_islands.pop_back();
    if (w.active == kNone) {
      w.active = static_cast<u32>(G.active_islands.size());
      G.active_islands.push_back(kept);
    }
  }
  l = Island{};
  G.free_handles.push_back(gone);
  return true;
}

// First publication: each island's node chain becomes its root followed
// by its other members in breadth-first discovery order from the
// component's lowest body id over the graph's edge lists
// (scene.cpp:5450-5626; with no lost edge yet, handles, roots and
// contact arrays are those the merges produced).
void Authority::rebuild_chains(Graph& G) {
  std::vector<uint8_t> done(n_, 0);
  std::vector<u32> members;
  for (u32 start = 0; start < n_; ++start) {
    if (!dynamic(start) || done[start]) continue;
    members.clear();
    members.push_back(start);
    done[start] = 1;
    for (std::size_t c = 0; c < members.size(); ++c) {
      for (u32 i = G.head[members[c]]; i != kNone;) {
        const Edge& e = edges_[i / 2];
        const u32 to = (i & 1) ? e.a : e.b;
        i = e.link[G.kind][i & 1].next;
        if (!dynamic(to) || done[to]) continue;
        done[to] = 1;
        members.push_back(to);
      }
    }
    Island& isl = G.islands[G.node_island[start]];
    const u32 root = isl.root;
    isl.tail = root;
    G.node_prev[root] = G.node_next[root] = kNone;
    for (u32 m : members) {
      if (m == root) continue;
      G.node_prev[m] = isl.tail;
      G.node_next[isl.tail] = m;
      G.node_next[m] = kNone;
      isl.tail = m;
    }
  }
}

// Binary min-heap on hop estimates with no tie break: a push stops at an
// equal parent, a pop prefers the left child on ties and stops before a
// child it is strictly below (scene.cpp:153-190, the order the route
// search visits nodes in).
void Authority::heap_push(Graph& G, u32 node) {
  auto& h = G.heap;
  const u64 entry = (u64(G.hops[node]) << 32) | node;
  std::size_t i = h.size();
  h.push_back(entry);
  while (i) {
    const std::size_t p = (i - 1) / 2;
    if (!((entry >> 32) < (h[p] >> 32))) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = entry;
}

u32 Authority::heap_pop(Graph& G) {
  auto& h = G.heap;
  const u32 top = static_cast<u32>(h.front());
  const u64 last = h.back();
  h.pop_back();
  const std::size_t n = h.size();
  std::size_t i = 0;
  for (;;) {
    std::size_t c = 2 * i + 1;
    if (c >= n) break;
    if (c + 1 < n && (h[c + 1] >> 32) < (h[c] >> 32)) ++c;
    if ((last >> 32) < (h[c] >> 32)) break;
    h[i] = h[c];
    i = c;
  }
  if (n) h[i] = last;
  return top;
}

bool Authority::live_hop(const Graph& G, u32 from, u32 to, u32 id) const {
  if (id == kNone || id >= edges_.size()) return false;
  const Edge& e = edges_[id];
  if (!e.live || e.retiring || (G.kind == 1 && !e.connected)) return false;
  if (!((e.a == from && e.b == to) || (e.a == to && e.b == from))) return false;
  return e.link[G.kind][0].linked && e.link[G.kind][1].linked;
}

// Lost-edge pass: each dirty node (ascending) keeps its island if a
// cached root route still holds or a hop-ordered search reaches the root
// or a node confirmed earlier in this pass; otherwise the searched
// component splits off (scene.cpp:4759-4982). Returns whether any split.
bool Authority::repair(Graph& G) {
  if (G.dirty.empty()) return false;
  std::sort(G.dirty.begin(), G.dirty.end());
  if (++G.confirm_stamp == 0) {
    for (auto& m : G.mark) m.confirmed = 0;
    G.confirm_stamp = 1;
  }
  const u32 cs = G.confirm_stamp;
  bool split_any = false;
  for (u32 node : G.dirty) {
    if (!dynamic(node) || G.mark[node].confirmed == cs) continue;
    const u32 isl = G.node_island[node];
    const u32 root = G.islands[isl].root;
    if (root == node) {
      G.mark[node].confirmed = cs;
      continue;
    }
    // cached route: strictly decreasing hops over live edges
    u32 probe = node;
    bool ok = false;
    G.members.clear();
    for (u32 depth = 0; depth < n_; ++depth) {
      if (probe == root || (G.mark[probe].confirmed == cs && G.node_island[probe] == isl)) {
        ok = true;
        break;
      }
      if (!G.hint[probe]) break;
      const u32 next = G.fast[probe];
      if (next == kNone || !dynamic(next) || !G.hint[next] || G.hops[next] >= G.hops[probe] ||
          !live_hop(G, probe, next, G.fast_edge[probe]))
        break;
      G.members.push_back(probe);
      probe = next;
    }
    if (ok) {
      for (u32 v : G.members) G.mark[v].confirmed = cs;
      continue;
    }
    // hop-ordered search; members are recorded when first reached
    if (++G.visit_stamp == 0) {
      for (auto& m : G.mark) m.seen = 0;
      G.visit_stamp = 1;
    }
    const u32 vs = G.visit_stamp;
    G.heap.clear();
    G.members.clear();
    G.members.push_back(node);
    heap_push(G, node);
    G.mark[node].seen = vs;
    G.mark[node].pred = kNone;
    G.mark[node].pred_edge = kNone;
    bool found = false;
    u32 witness = root;
    while (!G.heap.empty() && !found) {
      const u32 from = heap_pop(G);
      for (u32 i = G.head[from]; i != kNone;) {
        const Edge& e = edges_[i / 2];
        const u32 to = (i & 1) ? e.a : e.b;
        const u32 eid = i / 2;
        i = e.link[G.kind][i & 1].next;
        if (!dynamic(to) || G.mark[to].seen == vs) continue;
        G.mark[to].seen = vs;
        G.mark[to].pred = from;
        G.mark[to].pred_edge = eid;
        if (to == root || (G.mark[to].confirmed == cs && G.node_island[to] == isl)) {
          found = true;
          witness = to;
          break;
        }
        G.members.push_back(to);
        heap_push(G, to);
      }
    }
    if (!found) {
      split(G, node);
      for (u32 v : G.members) G.mark[v].confirmed = cs;
      split_any = true;
      continue;
    }
    if (witness == root) {
      G.fast[root] = kNone;
      G.fast_edge[root] = kNone;
      G.hops[root] = 0;
      G.hint[root] = 1;
    }
    for (u32 cur = witness; cur != node;) {
      const u32 p = G.mark[cur].pred;
      G.fast[p] = cur;
      G.fast_edge[p] = G.mark[cur].pred_edge;
      G.hops[p] = G.hops[cur] + 1;
      G.hint[p] = 1;
      G.mark[p].confirmed = cs;
      cur = p;
    }
    for (u32 v : G.members) {
      if (G.mark[v].confirmed == cs) continue;
      const u32 p = G.mark[v].pred;
      if (p == kNone) throw std::logic_error("route search lost a parent");
      G.fast[v] = p;
      G.fast_edge[v] = G.mark[v].pred_edge;
      G.hops[v] = G.hops[p] + 1;
      G.hint[v] = 1;
      G.mark[v].confirmed = cs;
    }
  }
  for (u32 v : G.dirty) G.dirty_mark[v] = 0;
  G.dirty.clear();
  return split_any;
}

// The searched component leaves its island: its contact managers are
// taken out of the old array in visit order (each once, at its first
// endpoint when that is dynamic), a handle comes off the free stack, the
// nodes keep their visit order as the new chain and the managers are
// appended in the order they were taken (scene.cpp:4984-5117).
void Authority::split(Graph& G, u32 seed) {
  const u32 owner = G.node_island[seed];
  const u32 root = G.islands[owner].root;
  if (root == seed || G.members.empty() || G.members.front() != seed ||
      G.members.size() >= G.islands[owner].count)
    throw std::logic_error("island split invariant lost");
  G.split_edges.clear();
  if (G.kind == 1) {
    for (u32 node : G.members)
      for (u32 i = G.head[node]; i != kNone;) {
        const u32 eid = i / 2;
        const Edge& e = edges_[eid];
        i = e.link[1][i & 1].next;
        if (!e.connected || (node != e.a && dynamic(e.a))) continue;
        remove_contact(eid);
        G.split_edges.push_back(eid);
      }
  }
  u32 created;
  if (!G.free_handles.empty()) {
    created = G.free_handles.back();
    G.free_handles.pop_back();
  } else {
    created = static_cast<u32>(G.islands.size());
    G.islands.emplace_back();
  }
  Island& old = G.islands[owner];
  for (u32 node : G.members) {
    const u32 before = G.node_prev[node], after = G.node_next[node];
    if (before != kNone) G.node_next[before] = after;
    else old.root = after;
    if (after != kNone) G.node_prev[after] = before;
    else old.tail = before;
    --old.count;
  }
  Island& isl = G.islands[created];
  isl = Island{};
  isl.live = true;
  isl.root = seed;
  isl.tail = G.members.back();
  isl.count = static_cast<u32>(G.members.size());
  if (G.islands[owner].active != kNone) {
    isl.active = static_cast<u32>(G.active_islands.size());
    G.active_islands.push_back(created);
  }
  for (std::size_t k = 0; k < G.members.size(); ++k) {
    const u32 node = G.members[k];
    G.node_prev[node] = k ? G.members[k - 1] : kNone;
    G.node_next[node] = k + 1 < G.members.size() ? G.members[k + 1] : kNone;
    G.node_island[node] = created;
    G.fast[node] = k ? G.mark[node].pred : kNone;
    G.fast_edge[node] = k ? G.mark[node].pred_edge : kNone;
    G.hops[node] = k ? G.hops[G.mark[node].pred] + 1 : 0;
    G.hint[node] = 1;
  }
  for (u32 eid : G.split_edges) append_contact(eid, created);
}

// Active nodes drop the bodies asleep (swap with the last; one inside the
// prefix active at the step's start first trades places with the
// prefix's last), the woken islands' chains are appended, islands with no
// active node leave the active list (swap with the last) and islands of
// active nodes not on it are appended in active-node order
// (scene.cpp:6751-6884).
void Authority::synchronize(Graph& G) {
  auto& an = G.active_nodes;
  for (std::size_t idx = 0; idx < an.size();) {
    const u32 id = an[idx];
    const bool keep = (G.kind == 0 || !sleeping_[id]) && G.node_island[id] != kNone;
    if (keep) {
      ++idx;
      continue;
    }
    if (idx < G.initial_active) {
      const std::size_t tail = G.initial_active - 1;
      const u32 prev = an[tail];
      an[idx] = prev;
      G.active_index[prev] = static_cast<u32>(idx);
      an[tail] = id;
      G.active_index[id] = static_cast<u32>(tail);
      --G.initial_active;
    }
    const u32 at = G.active_index[id];
    const u32 repl = an.back();
    an[at] = repl;
    G.active_index[repl] = at;
    an.pop_back();
    G.active_index[id] = kNone;
  }
  // an island's chain is appended once: a second pass over it would find
  // every member active already (or asleep and skipped again)
  if (G.expand_seen.size() < G.islands.size()) G.expand_seen.resize(G.islands.size(), 0);
  if (++G.expand_stamp == 0) {
    std::fill(G.expand_seen.begin(), G.expand_seen.end(), 0);
    G.expand_stamp = 1;
  }
  for (u32 w : G.activating) {
    const u32 h = G.node_island[w];
    if (h == kNone || !G.islands[h].live || G.expand_seen[h] == G.expand_stamp) continue;
    G.expand_seen[h] = G.expand_stamp;
    for (u32 id = G.islands[h].root; id != kNone; id = G.node_next[id]) {
      if ((G.kind == 1 && sleeping_[id]) || G.active_index[id] != kNone) continue;
      G.active_index[id] = static_cast<u32>(an.size());
      an.push_back(id);
    }
  }
  G.activating.clear();
  if (G.island_seen.size() < G.islands.size()) G.island_seen.resize(G.islands.size(), 0);
  if (++G.island_stamp == 0) {
    std::fill(G.island_seen.begin(), G.island_seen.end(), 0);
    G.island_stamp = 1;
  }
  const u32 st = G.island_stamp;
  for (u32 id : an) G.island_seen[G.node_island[id]] = st;
  auto& ai = G.active_islands;
  for (std::size_t idx = 0; idx < ai.size();) {
    const u32 h = ai[idx];
    if (G.islands[h].live && G.island_seen[h] == st) {
      G.islands[h].active = static_cast<u32>(idx);
      ++idx;
      continue;
    }
    G.islands[h].active = kNone;
    const u32 moved = ai.back();
    ai[idx] = moved;
    ai.pop_back();
    if (idx < ai.size()) G.islands[moved].active = static_cast<u32>(idx);
  }
  for (u32 id : an) {
    Island& isl = G.islands[G.node_island[id]];
    if (isl.active != kNone) continue;
    isl.active = static_cast<u32>(ai.size());
    ai.push_back(G.node_island[id]);
  }
}

// Wakes (scene.cpp:6148-6262): endpoints of new broadphase pairs with an
// awake side, then endpoints of new touches, queue for both graphs; every
// speculative island reached from the queue wakes its sleeping members,
// in chain order, and they join the accurate queue.
void Authority::activate(const fyzz_island_events& ev) {
  Graph& S = graph_[0];
  Graph& A = graph_[1];
  if (ev.touch_found_count == 0 && ev.pairs_found_count == 0) return;
  std::vector<uint8_t> qs(n_, 0), qa(n_, 0);
  for (u32 v : S.activating) qs[v] = 1;
  for (u32 v : A.activating) qa[v] = 1;
  auto push = [&](Graph& G, std::vector<uint8_t>& q, u32 v) {
    if (!dynamic(v) || q[v]) return;
    q[v] = 1;
    G.activating.push_back(v);
  };
  for (u32 k = 0; k < ev.pairs_found_count; ++k) {
    const u32 a = ev.pairs_found[2 * k], b = ev.pairs_found[2 * k + 1];
    if (!dynamic(a) || !dynamic(b)) continue;
    if (!sleeping_[a] || !sleeping_[b]) {
      push(S, qs, a);
      push(S, qs, b);
      push(A, qa, a);
      push(A, qa, b);
    }
  }
  for (u32 k = 0; k < ev.touch_found_count; ++k) {
    for (int s = 0; s < 2; ++s) {
      const u32 v = ev.touch_found[2 * k + s];
      if (!dynamic(v)) continue;
      push(A, qa, v);
      push(S, qs, v);
      ready_[v] = 0;
    }
  }
  if (S.island_seen.size() < S.islands.size()) S.island_seen.resize(S.islands.size(), 0);
  if (++S.island_stamp == 0) {
    std::fill(S.island_seen.begin(), S.island_seen.end(), 0);
    S.island_stamp = 1;
  }
  const u32 st = S.island_stamp;
  for (u32 w : S.activating) {
    const u32 h = S.node_island[w];
    if (h == kNone || !S.islands[h].live || S.island_seen[h] == st) continue;
    S.island_seen[h] = st;
    for (u32 v = S.islands[h].root; v != kNone; v = S.node_next[v]) {
      if (!sleeping_[v]) continue;
      push(A, qa, v);
      sleeping_[v] = 0;
      ready_[v] = 0;
    }
  }
}

// Sleeps (scene.cpp:6465-6536): active accurate islands, last first,
// whose members are all ready and one still awake, put their awake
// members to sleep in chain order.
void Authority::deactivate(Graph& G, std::vector<u32>& went) {
  went.clear();
  for (std::size_t k = G.active_islands.size(); k > 0; --k) {
    const Island& isl = G.islands[G.active_islands[k - 1]];
    if (!isl.live) continue;
    bool ready = true, awake = false;
    for (u32 v = isl.root; v != kNone; v = G.node_next[v]) {
      if (!ready_[v]) {
        ready = false;
        break;
      }
      awake |= !sleeping_[v];
    }
    if (!ready || !awake) continue;
    for (u32 v = isl.root; v != kNone; v = G.node_next[v]) {
      if (sleeping_[v]) continue;
      went.push_back(v);
      sleeping_[v] = 1;
    }
  }
}

static void check_sorted(const u32* p, u32 n, const char* what) {
  for (u32 k = 0; k < n; ++k) {
    if (p[2 * k] >= p[2 * k + 1]) throw std::invalid_argument(std::string(what) + ": pair not (low, high)");
    if (k && key(p[2 * k - 2], p[2 * k - 1]) >= key(p[2 * k], p[2 * k + 1]))
      throw std::invalid_argument(std::string(what) + ": pairs not ascending");
  }
}

void Authority::publish(const fyzz_island_events& ev, fyzz_island_order& out) {
  check_sorted(ev.pairs_found, ev.pairs_found_count, "pairs_found");
  check_sorted(ev.pairs_lost, ev.pairs_lost_count, "pairs_lost");
  check_sorted(ev.touch_found, ev.touch_found_count, "touch_found");
  check_sorted(ev.touch_lost, ev.touch_lost_count, "touch_lost");
  for (int g = 0; g < 2; ++g) graph_[g].initial_active = graph_[g].active_nodes.size();
  if (ev.ready)
    for (u32 v = 0; v < n_; ++v) ready_[v] = ev.ready[v] ? 1 : 0;

  // broadphase publication (scene.cpp:3233-3300): lost managers leave both
  // graphs and retire; found ones take a slot and join the speculative graph
  for (u32 k = 0; k < ev.pairs_lost_count; ++k) {
    const u32 id = edge_of(ev.pairs_lost[2 * k], ev.pairs_lost[2 * k + 1]);
    if (id == kNone) throw std::invalid_argument("pairs_lost: unknown pair");
    unlink(1, id);
    unlink(0, id);
    edges_[id].retiring = true;
    edges_[id].connected = false;
    by_pair_.erase(key(edges_[id].a, edges_[id].b));
    retired_edges_.push_back(id);
  }
  for (u32 k = 0; k < ev.pairs_found_count; ++k) {
    const u32 a = ev.pairs_found[2 * k], b = ev.pairs_found[2 * k + 1];
    if (a >= n_ || b >= n_) throw std::invalid_argument("pairs_found: body out of range");
    if (!dynamic(a) && !dynamic(b)) throw std::invalid_argument("pairs_found: two static bodies");
    if (edge_of(a, b) != kNone) throw std::invalid_argument("pairs_found: pair already present");
    u32 id;
    if (!free_edges_.empty()) {
      id = free_edges_.back();
      free_edges_.pop_back();
    } else {
      id = static_cast<u32>(edges_.size());
      edges_.emplace_back();
    }
    edges_[id] = Edge{};
    edges_[id].a = a;
    edges_[id].b = b;
    edges_[id].live = true;
    by_pair_.insert(key(a, b), id);
    link(0, id);
  }
  // speculative connectivity (scene.cpp:5841-5898)
  Graph& S = graph_[0];
  Graph& A = graph_[1];
  if (!admitted_) admit(S);
  for (u32 k = 0; k < ev.pairs_found_count; ++k) {
    const u32 a = ev.pairs_found[2 * k], b = ev.pairs_found[2 * k + 1];
    merge(S, a, b, edge_of(a, b));
  }
  if (!admitted_) rebuild_chains(S);

  // touch changes (scene.cpp:4702-4721): new touches link first, then the
  // lost ones unlink
  for (u32 k = 0; k < ev.touch_found_count; ++k) {
    const u32 id = edge_of(ev.touch_found[2 * k], ev.touch_found[2 * k + 1]);
    if (id == kNone) throw std::invalid_argument("touch_found: pair is not a broadphase pair");
    if (edges_[id].connected) throw std::invalid_argument("touch_found: pair already touching");
    link(1, id);
    edges_[id].connected = true;
  }
  for (u32 k = 0; k < ev.touch_lost_count; ++k) {
    const u32 id = edge_of(ev.touch_lost[2 * k], ev.touch_lost[2 * k + 1]);
    if (id == kNone) continue;  // its manager retired with the broadphase pair
    unlink(1, id);
    edges_[id].connected = false;
  }
  // accurate connectivity: merges, then the new touch joins its island's array
  if (!admitted_) admit(A);
  for (u32 k = 0; k < ev.touch_found_count; ++k) {
    const u32 a = ev.touch_found[2 * k], b = ev.touch_found[2 * k + 1];
    const u32 id = edge_of(a, b);
    merge(A, a, b, id);
    append_contact(id, A.node_island[owner_node(edges_[id])]);
  }
  if (!admitted_) rebuild_chains(A);
  admitted_ = true;

  activate(ev);
  // third pass: speculative then accurate route repair, active lists, sleep
  repair(S);
  synchronize(S);
  repair(A);
  synchronize(A);
  std::vector<u32> went;
  deactivate(A, went);
  if (!went.empty()) {
    synchronize(S);
    synchronize(A);
  }
  for (u32 id : retired_edges_) {
    edges_[id] = Edge{};
    free_edges_.push_back(id);
  }
  retired_edges_.clear();

  // solver order (scene.cpp:7174-7260)
  u32 m = 0, t = 0;
  std::size_t bodies = 0, rows = 0;
  for (u32 h : A.active_islands) {
    const Island& isl = A.islands[h];
    for (const Contact& c : isl.contacts) {
      if (m < out.manager_capacity) {
        out.managers[2 * m] = c.a;
        out.managers[2 * m + 1] = c.b;
      }
      ++m;
    }
    bodies += isl.count;
    rows += isl.contacts.size();
    if (bodies >= kTaskBodies && rows >= 1) {
      if (t < out.task_capacity) out.task_ends[t] = m;
      ++t;
      bodies = rows = 0;
    }
  }
  if (bodies) {
    if (t < out.task_capacity) out.task_ends[t] = m;
    ++t;
  }
  out.manager_count = m;
  out.task_count = t;
  if (out.sleeping)
    for (u32 v = 0; v < n_; ++v) out.sleeping[v] = sleeping_[v];
}

}  // namespace fyzz_hip::islands

struct fyzz_hip_islands {
  fyzz_hip::islands::Authority impl;
  std::string error;
  bool broken = false;  // a failed step may have applied part of its events
};

extern "C" {

int fyzz_hip_islands_create(uint32_t body_count, const uint8_t* dynamic, fyzz_hip_islands** out) {
  if (!out || !dynamic || body_count == 0 || body_count == 0xFFFFFFFFu) return FYZZ_HIP_E_INVALID_ARGUMENT;
  try {
    *out = new fyzz_hip_islands{fyzz_hip::islands::Authority(body_count, dynamic), {}};
    return FYZZ_HIP_OK;
  } catch (const std::bad_alloc&) {
    return FYZZ_HIP_E_OUT_OF_MEMORY;
  }
}

void fyzz_hip_islands_destroy(fyzz_hip_islands* s) { delete s; }

int fyzz_hip_islands_step(fyzz_hip_islands* s, const fyzz_island_events* events, fyzz_island_order* order) {
  if (!s || !events || !order) return FYZZ_HIP_E_INVALID_ARGUMENT;
  if ((order->manager_capacity && !order->managers) || (order->task_capacity && !order->task_ends))
    return FYZZ_HIP_E_INVALID_ARGUMENT;
  if (s->broken) return FYZZ_HIP_E_LOGIC;
  try {
    s->error.clear();
    s->impl.publish(*events, *order);
  } catch (const std::invalid_argument& e) {
    s->error = e.what();
    s->broken = true;
    return FYZZ_HIP_E_INVALID_ARGUMENT;
  } catch (const std::bad_alloc&) {
    s->error = "out of host memory";
    s->broken = true;
    return FYZZ_HIP_E_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    s->error = e.what();
    s->broken = true;
    return FYZZ_HIP_E_LOGIC;
  }
  if (order->manager_count > order->manager_capacity || order->task_count > order->task_capacity) {
    s->error = "order buffers too small";
    return FYZZ_HIP_E_OVERFLOW;
  }
  return FYZZ_HIP_OK;
}

const char* fyzz_hip_islands_last_error(const fyzz_hip_islands* s) { return s ? s->error.c_str() : ""; }

}  // extern "C"
