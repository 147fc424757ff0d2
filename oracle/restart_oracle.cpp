//========================================================================================
// restart_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of how the reference reads a restart file, the checker for the .rst
// writer in athena_amd/restart.py.  Only tests/ may load it.  It follows, for 3-D MHD
// without general relativity, user data, passive scalars or radiation:
//   ParameterInput::LoadFromFile / LoadFromStream / ParseLine
//     (parameter_input.cpp:120-180, 188-225, 265-298): the text up to "<par_end>",
//     blocks "<name>" and "name = value # comment" lines; the binary part starts 10
//     bytes after the start of "<par_end>";
//   the restart Mesh constructor (mesh/mesh.cpp:619-1010): a header of int nbtotal,
//     int root_level, RegionSize, Real time, Real dt, int ncycle and IOWrapperSizeT
//     datasize read as one packed block of 3 ints + 2 Reals + sizeof(RegionSize) +
//     sizeof(IOWrapperSizeT) bytes (mesh.cpp:710-738); the ID list of nbtotal
//     (LogicalLocation, Real cost) pairs (mesh.cpp:840-860); block size from
//     <meshblock> nx1..nx3 (default: the mesh size) (mesh.cpp:749-756); the block tree
//     rebuilt from the listed locations and the list replaced by the tree's own order
//     (tree.CreateRootGrid, AddMeshBlockWithoutRefine, GetMeshBlockList,
//     mesh.cpp:868-879); block i's data at headeroffset + i * datasize
//     (mesh.cpp:950-967), datasize checked against the block's size (mesh.cpp:983-989);
//   the restart MeshBlock constructor (meshblock.cpp:446-465): u, then b.x1f, b.x2f,
//     b.x3f, each an AthenaArray of (nx + 2 NGHOST) cells per direction (one more face
//     along its own direction).
// The tree order is restated here on its own: the root grid is the part of a complete
// octree of depth root_level that covers nrbx1 x nrbx2 x nrbx3 blocks, the listed finer
// locations are added along their paths, and the leaves are visited depth first with the
// children of a node in the order ox1 fastest, then ox2, then ox3 (meshblock_tree.cpp:
// 87-160, 336-356).  A file whose blocks are not in that order is read with the data of
// one block under the location of another, exactly as the reference would read it.
//========================================================================================
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

using Real = double;

struct RegionSize {            // athena.hpp:116-122 (x86-64 layout: 88 bytes)
  Real x1min, x2min, x3min;
  Real x1max, x2max, x3max;
  Real x1rat, x2rat, x3rat;
  int nx1, nx2, nx3;
};

struct LogicalLocation {       // athena.hpp:74-79 (32 bytes)
  std::int64_t lx1, lx2, lx3;
  int level;
};

using Params = std::map<std::string, std::map<std::string, std::string>>;

std::string g_err;

int fail(const std::string &m) {
  g_err = m;
  return 1;
}

// ParameterInput::LoadFromStream + ParseLine on the text before <par_end>.
void load_params(const std::string &text, Params &par) {
  std::istringstream is(text);
  std::string line, block;
  while (is.good()) {
    std::getline(is, line);
    std::string t;
    for (char ch : line)
      if (ch != '\t') t += ch;
    line = t;
    if (line.empty()) continue;
    std::size_t first = line.find_first_not_of(" ");
    if (first == std::string::npos) continue;
    if (line.compare(first, 1, "#") == 0) continue;
    if (line.compare(first, 9, "<par_end>") == 0) break;
    if (line.compare(first, 1, "<") == 0) {
      ++first;
      const std::size_t last = line.find_first_of(">", first);
      block.assign(line, first, last - 1);
      par[block];
      continue;
    }
    const std::size_t eq = line.find_first_of("=");
    const std::size_t hash = line.find_first_of("#");
    std::string name(line, first, eq - first);
    name.erase(name.find_last_not_of(" ") + 1, std::string::npos);
    std::string value(line, eq + 1, hash - eq - 1);
    value.erase(0, value.find_first_not_of(" "));
    value.erase(value.find_last_not_of(" ") + 1, std::string::npos);
    par[block][name] = value;
  }
}

int get_int(const Params &p, const char *b, const char *n, int dflt) {
  auto ib = p.find(b);
  if (ib == p.end()) return dflt;
  auto iv = ib->second.find(n);
  return (iv == ib->second.end()) ? dflt : std::stoi(iv->second);
}

// The block tree as the restart Mesh constructor rebuilds it (mesh.cpp:868-879):
// CreateRootGrid (meshblock_tree.cpp:87-110) makes the part of a complete octree of depth
// root_level that covers the root grid; AddMeshBlockWithoutRefine (:138-160) then walks
// from the root to each listed location, creating only the nodes on that path (a leaf it
// passes through gets its child array, with just that child, and stops being a leaf);
// GetMeshBlockList (:336-352) lists the leaves depth first, children n = i + 2j + 4k in
// increasing n.
struct TreeNode {
  LogicalLocation loc;
  int child[8];
  bool split;                  // the node has a child array (it is not a leaf)
};

struct Tree {
  std::vector<TreeNode> nodes;
  int add(const LogicalLocation &l) {
    TreeNode t;
    t.loc = l;
    for (int &c : t.child) c = -1;
    t.split = false;
    nodes.push_back(t);
    return static_cast<int>(nodes.size()) - 1;
  }
  void create_root_grid(int q, int root, const int *nrb) {
    const LogicalLocation l = nodes[q].loc;
    if (l.level == root) return;
    nodes[q].split = true;
    const std::int64_t levfac = std::int64_t(1) << (root - l.level - 1);
    for (int n = 0; n < 8; ++n) {
      const int i = n & 1, j = (n >> 1) & 1, k = (n >> 2) & 1;
      if ((l.lx3 * 2 + k) * levfac < nrb[2] && (l.lx2 * 2 + j) * levfac < nrb[1] &&
          (l.lx1 * 2 + i) * levfac < nrb[0]) {
        const int c = add({2 * l.lx1 + i, 2 * l.lx2 + j, 2 * l.lx3 + k, l.level + 1});
        nodes[q].child[n] = c;
        create_root_grid(c, root, nrb);
      }
    }
  }
  void add_without_refine(int q, const LogicalLocation &r) {
    const LogicalLocation l = nodes[q].loc;
    if (l.level == r.level) return;
    nodes[q].split = true;
    const int sh = r.level - l.level - 1;
    const int mx = ((r.lx1 >> sh) & 1) == 1, my = ((r.lx2 >> sh) & 1) == 1;
    const int mz = ((r.lx3 >> sh) & 1) == 1;
    const int n = mx + (my << 1) + (mz << 2);
    if (nodes[q].child[n] < 0) {
      const int c = add({2 * l.lx1 + mx, 2 * l.lx2 + my, 2 * l.lx3 + mz, l.level + 1});
      nodes[q].child[n] = c;
    }
    add_without_refine(nodes[q].child[n], r);
  }
  void list(int q, std::vector<LogicalLocation> &out) const {
    if (!nodes[q].split) {
      out.push_back(nodes[q].loc);
      return;
    }
    for (int n = 0; n < 8; ++n)
      if (nodes[q].child[n] >= 0) list(nodes[q].child[n], out);
  }
};

}  // namespace

extern "C" {

const char *ora_restart_error() { return g_err.c_str(); }

// Reads restart file `path` with nghost `ng`.  Writes ints[0..10] = nbtotal, root_level,
// ncycle, block nx1..nx3, mesh nx1..nx3, then the file's own list order flag (1 when
// the ID list equals the tree order); reals[0..10] = time, dt, x1min .. x3rat;
// locs[4 i ..] = (level, lx1, lx2, lx3) of block i in the order the reference assigns
// data to; data = every block's u, b.x1f, b.x2f, b.x3f in that order (cap doubles).
// Returns 0, or 1 with ora_restart_error() set.
int ora_restart_read(const char *path, int ng, int *ints, double *reals, long long *locs,
                     int cap_blocks, double *data, long cap_doubles) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail("cannot open the restart file");
  std::string raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  const std::size_t pe = raw.find("<par_end>", 0);
  if (pe == std::string::npos) return fail("<par_end> not found");
  Params par;
  load_params(raw.substr(0, pe + 10), par);
  std::size_t pos = pe + 10;                          // ParameterInput header length

  const std::size_t headersize = sizeof(int) * 3 + sizeof(Real) * 2 + sizeof(RegionSize) +
                                 sizeof(std::uint64_t);
  if (raw.size() < pos + headersize) return fail("the restart file is broken");
  const char *h = raw.data() + pos;
  int nbtotal, root_level, ncycle;
  RegionSize ms;
  Real time, dt;
  std::uint64_t datasize;
  std::size_t o = 0;
  std::memcpy(&nbtotal, h + o, sizeof(int)); o += sizeof(int);
  std::memcpy(&root_level, h + o, sizeof(int)); o += sizeof(int);
  std::memcpy(&ms, h + o, sizeof(RegionSize)); o += sizeof(RegionSize);
  std::memcpy(&time, h + o, sizeof(Real)); o += sizeof(Real);
  std::memcpy(&dt, h + o, sizeof(Real)); o += sizeof(Real);
  std::memcpy(&ncycle, h + o, sizeof(int)); o += sizeof(int);
  std::memcpy(&datasize, h + o, sizeof(std::uint64_t)); o += sizeof(std::uint64_t);
  pos += headersize;

  const int bnx[3] = {get_int(par, "meshblock", "nx1", ms.nx1),
                      get_int(par, "meshblock", "nx2", ms.nx2),
                      get_int(par, "meshblock", "nx3", ms.nx3)};
  const int nrb[3] = {ms.nx1 / bnx[0], ms.nx2 / bnx[1], ms.nx3 / bnx[2]};

  const std::size_t listsize = sizeof(LogicalLocation) + sizeof(Real);
  if (nbtotal < 1 || raw.size() < pos + listsize * nbtotal)
    return fail("the restart file is broken");
  std::vector<LogicalLocation> loclist(nbtotal);
  for (int i = 0; i < nbtotal; ++i) {
    std::memcpy(&loclist[i], raw.data() + pos + listsize * i, sizeof(LogicalLocation));
  }
  const std::size_t headeroffset = pos + listsize * nbtotal;

  // tree reconstruction and the list replaced by the tree's own order
  Tree bt;
  bt.add({0, 0, 0, 0});
  bt.create_root_grid(0, root_level, nrb);
  for (int i = 0; i < nbtotal; ++i) {
    if (loclist[i].level < root_level || loclist[i].level > root_level + 30)
      return fail("a listed location has an impossible level");
    bt.add_without_refine(0, loclist[i]);
  }
  std::vector<LogicalLocation> tree;
  bt.list(0, tree);
  if (static_cast<int>(tree.size()) != nbtotal)
    return fail("tree reconstruction failed: the total numbers of the blocks do not match");
  bool same_order = true;
  for (int i = 0; i < nbtotal; ++i)
    same_order = same_order && tree[i].lx1 == loclist[i].lx1 &&
                 tree[i].lx2 == loclist[i].lx2 && tree[i].lx3 == loclist[i].lx3 &&
                 tree[i].level == loclist[i].level;

  const long nc[3] = {bnx[0] + 2L * ng, bnx[1] + 2L * ng, bnx[2] + 2L * ng};
  const long nu = 5 * nc[2] * nc[1] * nc[0];
  const long nf1 = nc[2] * nc[1] * (nc[0] + 1), nf2 = nc[2] * (nc[1] + 1) * nc[0];
  const long nf3 = (nc[2] + 1) * nc[1] * nc[0];
  const long per = nu + nf1 + nf2 + nf3;
  if (datasize != static_cast<std::uint64_t>(per) * sizeof(Real))
    return fail("the restart file is broken or input parameters are inconsistent");
  if (raw.size() < headeroffset + datasize * nbtotal)
    return fail("the restart file is broken or input parameters are inconsistent");

  int *iv = ints;
  iv[0] = nbtotal; iv[1] = root_level; iv[2] = ncycle;
  iv[3] = bnx[0]; iv[4] = bnx[1]; iv[5] = bnx[2];
  iv[6] = ms.nx1; iv[7] = ms.nx2; iv[8] = ms.nx3;
  iv[9] = same_order ? 1 : 0;
  iv[10] = 0;
  const Real rv[11] = {time, dt, ms.x1min, ms.x2min, ms.x3min, ms.x1max, ms.x2max,
                       ms.x3max, ms.x1rat, ms.x2rat, ms.x3rat};
  for (int q = 0; q < 11; ++q) reals[q] = rv[q];
  if (cap_blocks >= nbtotal && locs) {
    for (int i = 0; i < nbtotal; ++i) {
      locs[4 * i] = tree[i].level;
      locs[4 * i + 1] = tree[i].lx1;
      locs[4 * i + 2] = tree[i].lx2;
      locs[4 * i + 3] = tree[i].lx3;
    }
  }
  if (data && cap_doubles >= per * nbtotal) {
    for (int i = 0; i < nbtotal; ++i)
      std::memcpy(data + per * i, raw.data() + headeroffset + datasize * i, datasize);
  }
  return 0;
}

}  // extern "C"
