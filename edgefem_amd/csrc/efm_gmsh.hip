// gmsh v2.2 ASCII reader (host code): the $Nodes and $Elements sections of
// load_gmsh_v2 (src/mesh_gmsh.cpp:15-101).  Nodes are `id x y z`; element lines are
// `id type ntags tags... nodes...`, of which type 2 (Tri3), 4 (Tet4) and 1 (Line2, kept as
// boundary lines) are read, with the first tag as the physical group; other types are
// skipped.  The reference parses with an iostream token loop on one thread; this reader
// loads the file once, splits each section at line boundaries into chunks parsed by
// worker threads (std::from_chars), and concatenates the chunks in file order, so every
// array is in the reference's order.  Edges are numbered afterwards, on the GPU
// (efm_topology_create).
#include "efm_internal.h"

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

struct efm_gmsh {
  std::vector<int64_t> node_ids;
  std::vector<double> xyz;
  std::vector<int64_t> tet_nodes, tet_elem, tri_nodes, tri_elem, line_nodes;
  std::vector<int32_t> tet_phys, tri_phys, line_phys;
};

namespace {

struct Cursor {
  const char *p, *end;
  void skip_blank() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r'))
      ++p;
  }
  template <typename T> bool num(T &v) {
    skip_blank();
    if (p < end && *p == '+')
      ++p;
    auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc())
      return false;
    p = r.ptr;
    return true;
  }
  void next_line() {
    const void *nl = std::memchr(p, '\n', (size_t)(end - p));
    p = nl ? static_cast<const char *>(nl) + 1 : end;
  }
};

// the section body between the line after `head` (the count line) and the `tail` marker
bool find_section(const std::string &text, const char *head, const char *tail, size_t &count, const char *&b,
                  const char *&e) {
  const size_t hl = std::strlen(head);
  size_t pos = 0;
  while (true) {
    pos = text.find(head, pos);
    if (pos == std::string::npos)
      return false;
    const bool at_start = pos == 0 || text[pos - 1] == '\n';
    const size_t after = pos + hl;
    const bool at_end = after >= text.size() || text[after] == '\n' || text[after] == '\r';
    if (at_start && at_end)
      break;
    pos = after;
  }
  Cursor c{text.data() + pos + hl, text.data() + text.size()};
  c.next_line();
  long long n = 0;
  if (!c.num(n) || n < 0)
    return false;
  count = (size_t)n;
  c.next_line();
  b = c.p;
  const size_t tp = text.find(tail, (size_t)(b - text.data()));
  e = tp == std::string::npos ? text.data() + text.size() : text.data() + tp;
  return true;
}

// [b, e) cut into k chunks at line starts
std::vector<const char *> split_lines(const char *b, const char *e, int k) {
  std::vector<const char *> cut{b};
  const size_t len = (size_t)(e - b);
  for (int i = 1; i < k; ++i) {
    const char *p = b + len * i / k;
    if (p <= cut.back())
      continue;
    const void *nl = std::memchr(p, '\n', (size_t)(e - p));
    p = nl ? static_cast<const char *>(nl) + 1 : e;
    if (p > cut.back() && p < e)
      cut.push_back(p);
  }
  cut.push_back(e);
  return cut;
}

struct NodeChunk {
  std::vector<int64_t> ids;
  std::vector<double> xyz;
  bool ok = true;
};

struct ElemChunk {
  std::vector<int64_t> tet, tet_e, tri, tri_e, lin;
  std::vector<int32_t> tet_p, tri_p, lin_p;
  size_t records = 0;  // element lines read, kept or skipped
  bool ok = true;
};

template <typename Chunk, typename Fn>
void parse_parallel(const std::vector<const char *> &cut, std::vector<Chunk> &out, Fn &&fn) {
  const int k = (int)cut.size() - 1;
  out.assign(k, Chunk{});
  std::vector<std::thread> th;
  for (int i = 1; i < k; ++i)
    th.emplace_back([&, i] { fn(Cursor{cut[i], cut[i + 1]}, out[i]); });
  if (k > 0)
    fn(Cursor{cut[0], cut[1]}, out[0]);
  for (auto &t : th)
    t.join();
}

}  // namespace

extern "C" int efm_gmsh_read(const char *path, int n_threads, efm_gmsh **out) {
  EFM_REQUIRE(path && out, "efm_gmsh_read: NULL argument");
  *out = nullptr;
  FILE *f = std::fopen(path, "rb");
  EFM_REQUIRE(f, std::string("Failed to open mesh file: ") + path);
  std::string text;
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz > 0) {
    text.resize((size_t)sz);
    const size_t got = std::fread(&text[0], 1, (size_t)sz, f);
    text.resize(got);
  }
  std::fclose(f);
  int k = n_threads > 0 ? n_threads : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  auto m = std::make_unique<efm_gmsh>();
  size_t nn = 0, ne = 0;
  const char *b, *e;
  if (find_section(text, "$Nodes", "$EndNodes", nn, b, e)) {
    std::vector<NodeChunk> ch;
    parse_parallel(split_lines(b, e, k), ch, [](Cursor c, NodeChunk &o) {
      while (true) {
        c.skip_blank();
        while (c.p < c.end && *c.p == '\n') {
          ++c.p;
          c.skip_blank();
        }
        if (c.p >= c.end)
          break;
        long long id;
        double x, y, z;
        if (!c.num(id) || !c.num(x) || !c.num(y) || !c.num(z)) {
          o.ok = false;
          return;
        }
        o.ids.push_back(id);
        o.xyz.insert(o.xyz.end(), {x, y, z});
        c.next_line();
      }
    });
    for (auto &c : ch) {
      EFM_REQUIRE(c.ok, "efm_gmsh_read: malformed $Nodes record");
      m->node_ids.insert(m->node_ids.end(), c.ids.begin(), c.ids.end());
      m->xyz.insert(m->xyz.end(), c.xyz.begin(), c.xyz.end());
    }
    EFM_REQUIRE(m->node_ids.size() == nn, "efm_gmsh_read: $Nodes count differs from the records");
  }
  if (find_section(text, "$Elements", "$EndElements", ne, b, e)) {
    std::vector<ElemChunk> ch;
    parse_parallel(split_lines(b, e, k), ch, [](Cursor c, ElemChunk &o) {
      while (true) {
        c.skip_blank();
        while (c.p < c.end && *c.p == '\n') {
          ++c.p;
          c.skip_blank();
        }
        if (c.p >= c.end)
          break;
        long long id;
        int type, ntag;
        if (!c.num(id) || !c.num(type) || !c.num(ntag) || ntag < 0) {
          o.ok = false;
          return;
        }
        int phys = 0;
        for (int t = 0; t < ntag; ++t) {
          int tag;
          if (!c.num(tag)) {
            o.ok = false;
            return;
          }
          if (t == 0)
            phys = tag;
        }
        const int nv = type == 4 ? 4 : type == 2 ? 3 : type == 1 ? 2 : 0;
        long long v[4];
        for (int q = 0; q < nv; ++q)
          if (!c.num(v[q])) {
            o.ok = false;
            return;
          }
        if (type == 4) {
          o.tet.insert(o.tet.end(), {v[0], v[1], v[2], v[3]});
          o.tet_p.push_back(phys);
          o.tet_e.push_back(id);
        } else if (type == 2) {
          o.tri.insert(o.tri.end(), {v[0], v[1], v[2]});
          o.tri_p.push_back(phys);
          o.tri_e.push_back(id);
        } else if (type == 1) {
          o.lin.insert(o.lin.end(), {v[0], v[1]});
          o.lin_p.push_back(phys);
        }
        ++o.records;
        c.next_line();
      }
    });
    size_t records = 0;
    for (auto &c : ch) {
      EFM_REQUIRE(c.ok, "efm_gmsh_read: malformed $Elements record");
      records += c.records;
      m->tet_nodes.insert(m->tet_nodes.end(), c.tet.begin(), c.tet.end());
      m->tet_phys.insert(m->tet_phys.end(), c.tet_p.begin(), c.tet_p.end());
      m->tet_elem.insert(m->tet_elem.end(), c.tet_e.begin(), c.tet_e.end());
      m->tri_nodes.insert(m->tri_nodes.end(), c.tri.begin(), c.tri.end());
      m->tri_phys.insert(m->tri_phys.end(), c.tri_p.begin(), c.tri_p.end());
      m->tri_elem.insert(m->tri_elem.end(), c.tri_e.begin(), c.tri_e.end());
      m->line_nodes.insert(m->line_nodes.end(), c.lin.begin(), c.lin.end());
      m->line_phys.insert(m->line_phys.end(), c.lin_p.begin(), c.lin_p.end());
    }
    EFM_REQUIRE(records == ne, "efm_gmsh_read: $Elements count differs from the records");
  }
  *out = m.release();
  return EFM_OK;
}

extern "C" int efm_gmsh_sizes(const efm_gmsh *m, int64_t *h_sizes) {
  EFM_REQUIRE(m && h_sizes, "efm_gmsh_sizes: NULL argument");
  h_sizes[0] = (int64_t)m->node_ids.size();
  h_sizes[1] = (int64_t)m->tet_phys.size();
  h_sizes[2] = (int64_t)m->tri_phys.size();
  h_sizes[3] = (int64_t)m->line_phys.size();
  return EFM_OK;
}

extern "C" int efm_gmsh_copy(const efm_gmsh *m, int64_t *node_ids, double *xyz, int64_t *tet_nodes,
                             int32_t *tet_phys, int64_t *tet_elem, int64_t *tri_nodes, int32_t *tri_phys,
                             int64_t *tri_elem, int64_t *line_nodes, int32_t *line_phys) {
  EFM_REQUIRE(m, "efm_gmsh_copy: NULL mesh");
  auto put = [](auto *dst, const auto &src) {
    if (dst && !src.empty())
      std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
  };
  put(node_ids, m->node_ids);
  put(xyz, m->xyz);
  put(tet_nodes, m->tet_nodes);
  put(tet_phys, m->tet_phys);
  put(tet_elem, m->tet_elem);
  put(tri_nodes, m->tri_nodes);
  put(tri_phys, m->tri_phys);
  put(tri_elem, m->tri_elem);
  put(line_nodes, m->line_nodes);
  put(line_phys, m->line_phys);
  return EFM_OK;
}

extern "C" void efm_gmsh_free(efm_gmsh *m) { delete m; }
