// Host symbolic phase: CRS pattern from the mesh, ILU(0) orderings, the device-numbering pattern and the
// phase/group/level schedules of the triangular sweeps and the numeric factorisation.
#include <algorithm>
#include <cstdlib>
#include <queue>
#include <cmath>
#include <deque>
#include <numeric>
#include <set>
#include <tuple>
#include <stdexcept>

#include "dfm_internal.hpp"

namespace dfm {

// CElemAry::MakePattern_FEM (src/femfield/elem_ary.cpp:225-245): node -> elements (MakePointSurElem,
// :675-724, ascending element index), then per node the distinct neighbours through its elements, self
// excluded (MakePointSurPoint with isnt_self, :821-892), ascending (CIndexedArray::Sort,
// include/delfem/indexed_array.h:113-132).
void build_pattern(int nnode, int nelem, int npoel, const int* lnods, std::vector<unsigned>& rowptr,
                   std::vector<unsigned>& colind, std::vector<int>& elsup_ptr, std::vector<int>& elsup) {
  elsup_ptr.assign((size_t)nnode + 1, 0);
  for (long long e = 0; e < nelem; e++)
    for (int k = 0; k < npoel; k++) elsup_ptr[lnods[e * npoel + k] + 1]++;
  for (int i = 0; i < nnode; i++) elsup_ptr[i + 1] += elsup_ptr[i];
  elsup.assign((size_t)elsup_ptr[nnode], 0);
  {
    std::vector<int> fill(elsup_ptr.begin(), elsup_ptr.end() - 1);
    for (long long e = 0; e < nelem; e++)
      for (int k = 0; k < npoel; k++) elsup[fill[lnods[e * npoel + k]]++] = (int)e;
  }
  rowptr.assign((size_t)nnode + 1, 0);
  std::vector<int> mark((size_t)nnode, -1);
  for (int i = 0; i < nnode; i++) {
    mark[i] = i;
    unsigned c = 0;
    for (int k = elsup_ptr[i]; k < elsup_ptr[i + 1]; k++) {
      const int* no = &lnods[(long long)elsup[k] * npoel];
      for (int a = 0; a < npoel; a++)
        if (mark[no[a]] != i) { mark[no[a]] = i; c++; }
    }
    rowptr[i + 1] = rowptr[i] + c;
  }
  colind.assign(rowptr[nnode], 0);
  std::fill(mark.begin(), mark.end(), -1);
  for (int i = 0; i < nnode; i++) {
    mark[i] = i;
    unsigned pos = rowptr[i];
    for (int k = elsup_ptr[i]; k < elsup_ptr[i + 1]; k++) {
      const int* no = &lnods[(long long)elsup[k] * npoel];
      for (int a = 0; a < npoel; a++)
        if (mark[no[a]] != i) { mark[no[a]] = i; colind[pos++] = (unsigned)no[a]; }
    }
    std::sort(colind.begin() + rowptr[i], colind.begin() + rowptr[i + 1]);
  }
}

// Permuted pattern as CMat_BlkCrs::AddPattern(rhs, order, order) builds it (src/matvec/mat_blkcrs.cpp:493-549:
// new row i0 takes old row order[i0], columns relabelled and sorted) and m_DiaInd
// (src/matvec/matdiafrac_blkcrs.cpp:263-277).
void build_device_pattern(Sys& s) {
  const int n = s.n;
  s.rowptr.assign((size_t)n + 1, 0);
  for (int i0 = 0; i0 < n; i0++) {
    const int i1 = s.order[i0];
    s.rowptr[i0 + 1] = s.rowptr[i0] + (s.rowptr0[i1 + 1] - s.rowptr0[i1]);
  }
  s.colind.assign(s.rowptr[n], 0);
  s.perm_crs.assign(s.rowptr0[n], 0);
  std::vector<std::pair<unsigned, unsigned>> tmp;
  for (int i0 = 0; i0 < n; i0++) {
    const int i1 = s.order[i0];
    tmp.clear();
    for (unsigned q = s.rowptr0[i1]; q < s.rowptr0[i1 + 1]; q++) tmp.push_back({(unsigned)s.inv[s.colind0[q]], q});
    std::sort(tmp.begin(), tmp.end());
    unsigned pos = s.rowptr[i0];
    for (auto& t : tmp) {
      s.colind[pos] = t.first;
      s.perm_crs[t.second] = pos;
      pos++;
    }
  }
  s.diaind.assign((size_t)n, 0);
  s.max_rowlen = 0;
  for (int i = 0; i < n; i++) {
    s.max_rowlen = std::max(s.max_rowlen, s.rowptr[i + 1] - s.rowptr[i]);
    s.diaind[i] = s.rowptr[i + 1];
    for (unsigned q = s.rowptr[i]; q < s.rowptr[i + 1]; q++)
      if (s.colind[q] > (unsigned)i) { s.diaind[i] = q; break; }
  }
  build_elem_lists(s);
}

// Element-store lists (Sys::eptr/eidx/eloc). A device row visits its incident elements in ascending index, so each
// entry's list comes out ascending.
void build_elem_lists(Sys& s) {
  const int n = s.n, NNO = s.npoel;
  s.eptr.clear();
  s.eidx.clear();
  if (s.lnods.empty() || (unsigned long long)s.nelem * NNO * NNO >= 0xFFFFFFFFull) return;
  if (s.eloc.size() != s.elsup.size()) {
    s.eloc.assign(s.elsup.size(), 0);
    for (int i1 = 0; i1 < n; i1++)
      for (int k = s.elsup_ptr[i1]; k < s.elsup_ptr[i1 + 1]; k++) {
        const int e = s.elsup[k];
        int a = 0;
        for (int c = 0; c < NNO; c++)
          if (s.lnods[(size_t)e * NNO + c] == i1) a = c;
        s.eloc[k] = (unsigned)e * NNO + (unsigned)a;
      }
  }
  const size_t ncrs = s.colind.size();
  s.eptr.assign(ncrs + 1, 0);
  auto entry = [&](int i0, unsigned j0) {
    const unsigned* b = s.colind.data() + s.rowptr[i0];
    const unsigned* e = s.colind.data() + s.rowptr[i0 + 1];
    return (size_t)(std::lower_bound(b, e, j0) - s.colind.data());
  };
  for (int pass = 0; pass < 2; pass++) {
    std::vector<unsigned> fill;
    if (pass == 1) {
      for (size_t q = 0; q < ncrs; q++) s.eptr[q + 1] += s.eptr[q];
      s.eidx.assign(s.eptr[ncrs], 0);
      fill.assign(s.eptr.begin(), s.eptr.end() - 1);
    }
    for (int i0 = 0; i0 < n; i0++) {
      const int i1 = s.order[i0];
      for (int k = s.elsup_ptr[i1]; k < s.elsup_ptr[i1 + 1]; k++) {
        const unsigned ea = s.eloc[k], e = ea / NNO, a = ea % NNO;
        for (int c = 0; c < NNO; c++) {
          if ((unsigned)c == a) continue;
          const size_t q = entry(i0, (unsigned)s.inv[s.lnods[(size_t)e * NNO + c]]);
          if (pass == 0) s.eptr[q + 1]++;
          else s.eidx[fill[q]++] = ea * NNO + (unsigned)c;
        }
      }
    }
  }
}

static const int kGroup = 64;  // rows per group: one wavefront, one row per lane

// Spatial key of each device row: its node's coordinate along the longest side of the mesh's bounding box
// (empty when the system has no coordinates).
static std::vector<double> row_key(const Sys& s) {
  std::vector<double> key;
  if (s.ndim < 1 || s.coords.size() != (size_t)s.n * s.ndim || s.order.size() != (size_t)s.n) return key;
  if (const char* e = std::getenv("DFM_XCD_PLACEMENT"); e && e[0] == '0') return key;
  int axis = 0;
  double best = -1.0;
  for (int d = 0; d < s.ndim; d++) {
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < s.n; i++) {
      lo = std::min(lo, s.coords[(size_t)i * s.ndim + d]);
      hi = std::max(hi, s.coords[(size_t)i * s.ndim + d]);
    }
    if (hi - lo >= best) { best = hi - lo; axis = d; }
  }
  key.resize((size_t)s.n);
  for (int r = 0; r < s.n; r++) key[r] = s.coords[(size_t)s.order[r] * s.ndim + axis];
  return key;
}

// The dispatcher hands the workgroups of a launch to the eight XCDs in turn (workgroup b to XCD b % 8), and every
// XCD has its own L2. xcd_positions orders m independent work items (item k has spatial key key[k]) over npos >= m
// positions, per_wg consecutive positions per workgroup, so that each XCD receives one spatial strip: the items
// sorted by key are cut into eight runs, run x filling XCD x's positions in ascending item order. Neighbouring
// items then read the lines of their shared neighbours through one L2 instead of eight. With npos == m every
// position is filled (runs sized to the XCDs' shares); otherwise the runs are balanced and the spare positions
// hold -1.
static constexpr int kXcd = 8;

static std::vector<int> xcd_positions(const std::vector<double>& key, int npos, int per_wg) {
  const int m = (int)key.size();
  std::vector<int> sorted((size_t)m);
  std::iota(sorted.begin(), sorted.end(), 0);
  std::stable_sort(sorted.begin(), sorted.end(), [&](int a, int b) { return key[a] < key[b]; });
  std::vector<long long> cap(kXcd, 0), run(kXcd, 0);
  for (int i = 0; i < npos; i++) cap[(i / per_wg) % kXcd]++;
  for (int x = 0; x < kXcd; x++) run[x] = npos == m ? cap[x] : (long long)m * (x + 1) / kXcd - (long long)m * x / kXcd;
  std::vector<long long> next(kXcd, 0), end(kXcd, 0);
  long long acc = 0;
  for (int x = 0; x < kXcd; x++) {
    if (run[x] > cap[x]) throw std::runtime_error("schedule: XCD run larger than its positions");
    next[x] = acc;
    acc += run[x];
    end[x] = acc;
    // inside a run the items keep their own order (for slices and tiles: device order, so each XCD streams its
    // strip's values through ascending addresses)
    std::sort(sorted.begin() + next[x], sorted.begin() + end[x]);
  }
  std::vector<int> pos((size_t)npos, -1);
  for (int i = 0; i < npos; i++) {
    const int x = (i / per_wg) % kXcd;
    if (next[x] < end[x]) pos[i] = sorted[(size_t)next[x]++];
  }
  return pos;
}

// Mean key of the rows [a, b) (a group or a slice).
static double range_key(const std::vector<double>& key, int a, int b) {
  double t = 0.0;
  for (int r = a; r < b; r++) t += key[r];
  return b > a ? t / (b - a) : 0.0;
}

// ILU factor pattern in the device numbering. ILU(0): the matrix pattern. ILU(k): the level-of-fill pattern of
// CMatDiaFrac_BlkCrs::AddFracPtn (src/matvec/matdiafrac_blkcrs.cpp:1254-1381), row by row: the row's entries
// with level < k pivot in ascending column order, each spreading the upper entries of its own (final) factor
// row whose level is < k; an entry met for the first time gets level max(lev_ik, lev_kj) + 1 (:1354-1356) and
// keeps it (a later, lower level does not replace it). Pivots are taken from a min-heap, so fill below the
// diagonal pivots in its turn. Then the matrix entry -> factor entry map for the value scatter (SetValue).
static void build_factor_pattern(Sys& s) {
  const int n = s.n;
  if (s.fill_level == 0) {
    s.frowptr = s.rowptr;
    s.fcolind = s.colind;
    s.fdiaind = s.diaind;
    s.a2f.clear();
    return;
  }
  const int kmax = s.fill_level;  // -1: no limit
  auto pivots = [&](int lev) { return kmax < 0 || lev < kmax; };
  std::vector<unsigned> fp((size_t)n + 1, 0), fc, fd((size_t)n, 0);
  std::vector<int> fl;  // level of every factor entry
  fc.reserve(s.colind.size() * 2);
  fl.reserve(s.colind.size() * 2);
  std::vector<int> lev_of((size_t)n, -1);  // level of column j in the current row, -1: absent
  std::vector<int> touched;
  std::priority_queue<int, std::vector<int>, std::greater<int>> heap;
  for (int i = 0; i < n; i++) {
    touched.clear();
    for (unsigned q = s.rowptr[i]; q < s.rowptr[i + 1]; q++) {
      const int j = (int)s.colind[q];
      lev_of[j] = 0;
      touched.push_back(j);
      if (j < i && pivots(0)) heap.push(j);
    }
    while (!heap.empty()) {
      const int k = heap.top();
      heap.pop();
      const int lik = lev_of[k];
      for (unsigned kj = fd[k]; kj < fp[k + 1]; kj++) {
        const int lkj = fl[kj];
        if (!pivots(lkj)) continue;
        const int j = (int)fc[kj];
        if (j == i || lev_of[j] >= 0) continue;
        const int l = std::max(lik, lkj) + 1;
        lev_of[j] = l;
        touched.push_back(j);
        if (j < i && pivots(l)) heap.push(j);
      }
    }
    std::sort(touched.begin(), touched.end());
    fd[i] = fp[i] + (unsigned)touched.size();
    for (size_t t = 0; t < touched.size(); t++) {
      const int j = touched[t];
      if (j > i && fd[i] == fp[i] + touched.size()) fd[i] = fp[i] + (unsigned)t;
      fc.push_back((unsigned)j);
      fl.push_back(lev_of[j]);
      lev_of[j] = -1;
    }
    if (fc.size() > 0xFFFFFFF0ull) throw std::runtime_error("ILU(k): factor pattern too large");
    fp[i + 1] = (unsigned)fc.size();
  }
  s.frowptr.swap(fp);
  s.fcolind.swap(fc);
  s.fdiaind.swap(fd);
  // both rows sorted: merge
  s.a2f.assign(s.colind.size(), 0);
  for (int i = 0; i < n; i++) {
    unsigned f = s.frowptr[i];
    for (unsigned q = s.rowptr[i]; q < s.rowptr[i + 1]; q++) {
      while (s.fcolind[f] != s.colind[q]) f++;
      s.a2f[q] = f;
    }
  }
}

// Level sets of the whole matrix (natural or arbitrary orderings): every level is a phase of independent
// rows, handed out in chunks of 64 through a row list.
static void level_sweeps(Sys& s) {
  const int n = s.n;
  for (int dir = 0; dir < 2; dir++) {
    Sweep& w = dir == 0 ? s.fwd : s.bwd;
    std::vector<int> lev((size_t)n, 0);
    int nlev = 0;
    if (dir == 0) {
      for (int i = 0; i < n; i++) {
        int l = 0;
        for (unsigned q = s.frowptr[i]; q < s.fdiaind[i]; q++) l = std::max(l, lev[s.fcolind[q]] + 1);
        lev[i] = l;
        nlev = std::max(nlev, l + 1);
      }
    } else {
      for (int i = n - 1; i >= 0; i--) {
        int l = 0;
        for (unsigned q = s.fdiaind[i]; q < s.frowptr[i + 1]; q++) l = std::max(l, lev[s.fcolind[q]] + 1);
        lev[i] = l;
        nlev = std::max(nlev, l + 1);
      }
    }
    std::vector<int> cnt((size_t)nlev + 1, 0);
    for (int i = 0; i < n; i++) cnt[lev[i] + 1]++;
    for (int l = 0; l < nlev; l++) cnt[l + 1] += cnt[l];
    w.rowlist.assign((size_t)n, 0);
    {
      std::vector<int> fill(cnt.begin(), cnt.end() - 1);
      for (int i = 0; i < n; i++) w.rowlist[fill[lev[i]]++] = i;
    }
    // groups are (start, end) pairs into the row list, in execution order
    w.phase_gptr.assign(1, 0);
    w.phase_nlev.clear();
    w.grp.clear();
    for (int l = 0; l < nlev; l++) {
      for (int r = cnt[l]; r < cnt[l + 1]; r += kGroup) {
        w.grp.push_back(r);
        w.grp.push_back(std::min(r + kGroup, cnt[l + 1]));
      }
      w.phase_gptr.push_back((int)w.grp.size() / 2);
      w.phase_nlev.push_back(1);
    }
    w.lev.assign((size_t)n, 0);
  }
  s.lsplit = s.fdiaind;
  s.usplit = s.fdiaind;
}

// Contiguous groups (MC / BMC orderings): group g = device rows [gstart[g], gstart[g+1]); dependencies inside a
// group are resolved by intra-group levels, those on other groups by phases: a group's forward phase is one more
// than the latest phase of the groups its rows' lower entries reach (backward: upper entries), so every group
// runs after the groups it reads. For ILU(0) of a colour-major ordering this is at most the colour; fill (ILU(k))
// may link groups of one colour, which then take later phases.
static void group_sweeps(Sys& s, const std::vector<int>& gstart) {
  const int n = s.n;
  const std::vector<double> key = row_key(s);
  const int ngroup = (int)gstart.size() - 1;
  const std::vector<unsigned>& rp = s.frowptr;
  const std::vector<unsigned>& ci = s.fcolind;
  const std::vector<unsigned>& di = s.fdiaind;
  std::vector<int> grp_of((size_t)n);
  for (int g = 0; g < ngroup; g++)
    for (int r = gstart[g]; r < gstart[g + 1]; r++) grp_of[r] = g;
  s.lsplit.assign((size_t)n, 0);
  s.usplit.assign((size_t)n, 0);
  std::vector<int> flev((size_t)n, 0), blev((size_t)n, 0), fph((size_t)ngroup, 0), bph((size_t)ngroup, 0);
  for (int i = 0; i < n; i++) {
    const int g = grp_of[i], g0 = gstart[g];
    unsigned ls = di[i];
    for (unsigned q = rp[i]; q < di[i]; q++) {
      const int j = (int)ci[q];
      if (j >= g0) { ls = q; break; }
      fph[g] = std::max(fph[g], fph[grp_of[j]] + 1);
    }
    s.lsplit[i] = ls;
    int l = 0;
    for (unsigned q = ls; q < di[i]; q++) l = std::max(l, flev[ci[q]] + 1);
    flev[i] = l;
  }
  for (int i = n - 1; i >= 0; i--) {
    const int g = grp_of[i], g1 = gstart[g + 1];
    unsigned us = rp[i + 1];
    for (unsigned q = di[i]; q < rp[i + 1]; q++) {
      const int j = (int)ci[q];
      if (j >= g1) {
        if (us == rp[i + 1]) us = q;
        bph[g] = std::max(bph[g], bph[grp_of[j]] + 1);
      }
    }
    s.usplit[i] = us;
    int l = 0;
    for (unsigned q = di[i]; q < us; q++) l = std::max(l, blev[ci[q]] + 1);
    blev[i] = l;
  }
  for (int dir = 0; dir < 2; dir++) {
    Sweep& w = dir == 0 ? s.fwd : s.bwd;
    const std::vector<int>& lev = dir == 0 ? flev : blev;
    const std::vector<int>& ph = dir == 0 ? fph : bph;
    int nph = 0;
    for (int g = 0; g < ngroup; g++) nph = std::max(nph, ph[g] + 1);
    w.rowlist.clear();
    w.lev.assign((size_t)n, 0);
    for (int i = 0; i < n; i++) {
      if (lev[i] > 255) throw std::runtime_error("schedule: more than 256 levels inside a group");
      w.lev[i] = (uint8_t)lev[i];
    }
    // groups of each phase (independent of one another) in device order, then placed over the XCDs by
    // position; (start, end) pairs in execution order
    std::vector<std::vector<int>> members((size_t)nph);
    for (int g = 0; g < ngroup; g++) members[ph[g]].push_back(g);
    if (!key.empty())
      for (auto& mem : members) {
        std::vector<double> gk(mem.size());
        for (size_t k = 0; k < mem.size(); k++) gk[k] = range_key(key, gstart[mem[k]], gstart[mem[k] + 1]);
        const std::vector<int> pos = xcd_positions(gk, (int)mem.size(), kSweepWavesPerWG);
        std::vector<int> placed(mem.size());
        for (size_t k = 0; k < mem.size(); k++) placed[k] = mem[pos[k]];
        mem.swap(placed);
      }
    w.grp.clear();
    w.phase_gptr.assign(1, 0);
    w.phase_nlev.clear();
    for (int p = 0; p < nph; p++) {
      int nl = 1;
      for (int g : members[p]) {
        for (int r = gstart[g]; r < gstart[g + 1]; r++) nl = std::max(nl, lev[r] + 1);
        w.grp.push_back(gstart[g]);
        w.grp.push_back(gstart[g + 1]);
      }
      w.phase_gptr.push_back((int)w.grp.size() / 2);
      w.phase_nlev.push_back(nl);
    }
  }
}

// Packed layout of one sweep (see Sweep): per group, the entries of each row split into "outside" (rows of
// earlier phases, final before the phase starts) and "in-group" (resolved level by level inside the wavefront),
// each class in ascending column order as the reference's substitution loops visit them. A class is stored as
// jagged rows k < 4 (row k: the k-th entry of every lane that has one, in lane order), then the entries k >= 4
// lane after lane; no padding.
static constexpr int kJagRows = 4;  // the device's kJag

static void build_slabs(const Sys& s, Sweep& w, bool forward) {
  const int ngroup = (int)w.grp.size() / 2;
  const bool indirect = !w.rowlist.empty();
  w.gmeta.assign((size_t)4 * ngroup, 0);
  w.srow.assign((size_t)kGroup * ngroup, -1);
  w.scnt.assign((size_t)kGroup * ngroup, 0u);
  w.ecol.clear();
  w.esrc.clear();
  long long off = 0;
  for (int g = 0; g < ngroup; g++) {
    const int a = w.grp[2 * g], e = w.grp[2 * g + 1], nrows = e - a;
    if (nrows < 1 || nrows > kGroup) throw std::runtime_error("schedule: bad group size");
    int wo = 0, wi = 0, nlev = 1;
    // per lane: [o0, o1) outside entries, [i0, i1) in-group entries
    unsigned o0[kGroup], o1[kGroup], i0[kGroup], i1[kGroup];
    long long nent = 0, nout = 0;
    for (int l = 0; l < nrows; l++) {
      const int row = indirect ? w.rowlist[a + l] : a + l;
      if (forward) {
        o0[l] = s.frowptr[row]; o1[l] = s.lsplit[row]; i0[l] = s.lsplit[row]; i1[l] = s.fdiaind[row];
      } else {
        i0[l] = s.fdiaind[row]; i1[l] = s.usplit[row]; o0[l] = s.usplit[row]; o1[l] = s.frowptr[row + 1];
      }
      const int no = (int)(o1[l] - o0[l]), ni = (int)(i1[l] - i0[l]);
      if (no > 255 || ni > 255) throw std::runtime_error("schedule: more than 255 entries per row and class");
      if (ni && indirect) throw std::runtime_error("schedule: in-group entries in an indirect group");
      wo = std::max(wo, no);
      wi = std::max(wi, ni);
      nent += no + ni;
      nout += no;
      nlev = std::max(nlev, (int)w.lev[row] + 1);
      w.srow[(size_t)g * kGroup + l] = row;
      w.scnt[(size_t)g * kGroup + l] = (unsigned)no | ((unsigned)ni << 8) | ((unsigned)w.lev[row] << 16);
    }
    if (off + nent > INT32_MAX) throw std::runtime_error("schedule: factor too large for 32-bit entry offsets");
    w.gmeta[4 * g + 0] = (int)off;
    w.gmeta[4 * g + 1] = wo | (wi << 8) | (nlev << 16);
    w.gmeta[4 * g + 2] = indirect ? -1 : a;
    w.gmeta[4 * g + 3] = nrows | (int)(nout << 8);
    // one class: jagged rows, then the rest lane by lane
    auto put = [&](const unsigned* c0, const unsigned* c1, int width, bool ingroup) {
      auto push = [&](unsigned q) {
        w.ecol.push_back(ingroup ? s.fcolind[q] - (unsigned)a : s.fcolind[q]);  // in-group: lane of the source row
        w.esrc.push_back(q);
      };
      for (int k = 0; k < std::min(width, kJagRows); k++)
        for (int l = 0; l < nrows; l++)
          if (c0[l] + k < c1[l]) push(c0[l] + k);
      for (int l = 0; l < nrows; l++)
        for (unsigned q = c0[l] + kJagRows; q < c1[l]; q++) push(q);
    };
    put(o0, o1, wo, false);
    put(i0, i1, wi, true);
    off += nent;
  }
  w.nent = off;
  w.phase_wo.assign(w.phase_nlev.size(), 0);
  w.phase_wi.assign(w.phase_nlev.size(), 0);
  for (size_t p = 0; p < w.phase_wo.size(); p++)
    for (int g = w.phase_gptr[p]; g < w.phase_gptr[p + 1]; g++) {
      w.phase_wo[p] = std::max(w.phase_wo[p], w.gmeta[4 * g + 1] & 255);
      w.phase_wi[p] = std::max(w.phase_wi[p], (w.gmeta[4 * g + 1] >> 8) & 255);
    }
}

// SELL-64 layout of the device-numbering pattern (see Sell).
static void build_sell(Sys& s) {
  Sell& c = s.sell;
  const int n = s.n, ns = (n + kGroup - 1) / kGroup;
  c.sptr.assign((size_t)ns + 1, 0);
  for (int sl = 0; sl < ns; sl++) {
    unsigned w = 0;
    for (int r = sl * kGroup; r < std::min(n, (sl + 1) * kGroup); r++) w = std::max(w, s.rowptr[r + 1] - s.rowptr[r]);
    c.sptr[sl + 1] = c.sptr[sl] + (int)w;
  }
  c.nslab = c.sptr[ns];
  // visiting order of the slices over the XCDs: positions of kSellSlicesPerWG slices per workgroup, the
  // workgroup count rounded up to a multiple of the XCD count
  const int nwg = ((ns + kSellSlicesPerWG - 1) / kSellSlicesPerWG + kXcd - 1) / kXcd * kXcd;
  const std::vector<double> key = row_key(s);
  if (key.empty() || ns == 0) {
    c.visit.resize((size_t)nwg * kSellSlicesPerWG);
    for (size_t i = 0; i < c.visit.size(); i++) c.visit[i] = i < (size_t)ns ? (int)i : -1;
  } else {
    std::vector<double> sk((size_t)ns);
    for (int sl = 0; sl < ns; sl++) sk[sl] = range_key(key, sl * kGroup, std::min(n, (sl + 1) * kGroup));
    c.visit = xcd_positions(sk, nwg * kSellSlicesPerWG, kSellSlicesPerWG);
  }
  c.col.assign((size_t)c.nslab * kGroup, 0u);
  c.src.assign((size_t)c.nslab * kGroup, ~0u);
  for (int r = 0; r < n; r++) {
    const int sl = r / kGroup, l = r % kGroup;
    for (unsigned q = s.rowptr[r]; q < s.rowptr[r + 1]; q++) {
      const size_t k = ((size_t)c.sptr[sl] + (q - s.rowptr[r])) * kGroup + l;
      c.col[k] = s.colind[q];
      c.src[k] = q;
    }
  }
}

// The last forward phase and the first backward phase may hold the same tiles (ILU(0) of a colour-major
// ordering: the last colour, which no later row needs before the backward sweep and which needs no later row):
// then one wavefront can run a tile's forward sweep and, with the values still in registers, its backward sweep.
static void fuse_last_first(Sys& s) {
  Sweep& f = s.fwd;
  const Sweep& b = s.bwd;
  f.fuse_bgrp.clear();
  if (f.nphase() == 0 || b.nphase() == 0) return;
  const int f0 = f.phase_gptr[f.nphase() - 1], f1 = f.phase_gptr[f.nphase()];
  const int b0 = b.phase_gptr[0], b1 = b.phase_gptr[1];
  if (f1 - f0 != b1 - b0) return;
  std::vector<std::pair<int, int>> byrow;  // (first row, backward group)
  for (int g = b0; g < b1; g++) {
    if (b.grp[2 * g + 1] - b.grp[2 * g] < 1) return;
    byrow.push_back({b.grp[2 * g], g});
  }
  std::sort(byrow.begin(), byrow.end());
  std::vector<int> map;
  for (int g = f0; g < f1; g++) {
    auto it = std::lower_bound(byrow.begin(), byrow.end(), std::make_pair(f.grp[2 * g], -1));
    if (it == byrow.end() || it->first != f.grp[2 * g]) return;
    const int bg = it->second;
    if (b.grp[2 * bg + 1] != f.grp[2 * g + 1]) return;
    map.push_back(bg);
  }
  f.fuse_bgrp.swap(map);
}

void build_schedule(Sys& s) {
  build_sell(s);
  build_sweeps(s);
}

// The lane plan of the numeric factorisation over the forward schedule (FactorPlan). Off (the per-row kernel
// runs instead) when a row is wider than a wavefront or the update lists outgrow 32-bit offsets.
static void build_factor_plan(Sys& s) {
  FactorPlan& fp = s.fplan;
  const int n = s.n;
  const std::vector<unsigned>& rp = s.frowptr;
  const std::vector<unsigned>& ci = s.fcolind;
  const std::vector<unsigned>& di = s.fdiaind;
  fp.ok = false;
  fp.gpass.clear(); fp.pmeta.clear(); fp.etab.clear(); fp.utab.clear(); fp.ukj.clear(); fp.ustep.clear();
  for (int i = 0; i < n; i++)
    if (rp[i + 1] - rp[i] + 1 > 64) return;
  fp.ukj.reserve(rp[n] * 2);
  fp.ustep.reserve(rp[n] * 2);
  std::vector<int> pos((size_t)n, -1);
  std::vector<std::vector<unsigned>> lkj(64);
  std::vector<std::vector<uint8_t>> lst(64);
  // a row's lanes: lower entries, the diagonal, upper entries; their update lists in pass order
  auto emit_row = [&](int i) {
    const unsigned r0 = rp[i], r1 = rp[i + 1], d = di[i];
    const int nl = (int)(d - r0), w = (int)(r1 - r0) + 1;
    for (int t = 0; t < w; t++) lkj[t].clear(), lst[t].clear();
    for (unsigned q = r0; q < r1; q++) pos[ci[q]] = (int)(q - r0) + (q >= d ? 1 : 0);
    pos[i] = nl;
    for (unsigned ik = r0; ik < d; ik++) {
      const unsigned k = ci[ik];
      for (unsigned kj = di[k]; kj < rp[k + 1]; kj++) {
        const int t = pos[ci[kj]];
        if (t >= 0) lkj[t].push_back(kj), lst[t].push_back((uint8_t)(ik - r0));
      }
    }
    for (unsigned q = r0; q < r1; q++) pos[ci[q]] = -1;
    pos[i] = -1;
    for (int t = 0; t < w; t++) {
      fp.etab.push_back(t == nl ? (unsigned)i : r0 + (unsigned)(t < nl ? t : t - 1));
      fp.utab.push_back((unsigned)fp.ukj.size());
      fp.ukj.insert(fp.ukj.end(), lkj[t].begin(), lkj[t].end());
      fp.ustep.insert(fp.ustep.end(), lst[t].begin(), lst[t].end());
    }
  };
  const Sweep& w = s.fwd;
  const int ngroup = (int)(w.grp.size() / 2);
  fp.gpass.assign(1, 0);
  std::vector<int> rows;
  for (int g = 0; g < ngroup; g++) {
    rows.clear();
    for (int r = w.grp[2 * g]; r < w.grp[2 * g + 1]; r++) rows.push_back(w.rowlist.empty() ? r : w.rowlist[r]);
    std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) { return w.lev[a] < w.lev[b]; });
    size_t k = 0;
    int prev_lev = -1;
    while (k < rows.size()) {
      const int lev = w.lev[rows[k]];
      unsigned long long smask = 0, dmask = 0;
      int lanes = 0, steps = 0;
      while (k < rows.size() && w.lev[rows[k]] == lev) {
        const int i = rows[k];
        const int wi = (int)(rp[i + 1] - rp[i]) + 1, nl = (int)(di[i] - rp[i]);
        if (lanes + wi > 64) break;
        smask |= 1ull << lanes;
        dmask |= 1ull << (lanes + nl);
        lanes += wi;
        steps = std::max(steps, nl);
        emit_row(i);
        k++;
      }
      for (int l = lanes; l < 64; l++) {  // idle lanes: no entry, an empty update range
        fp.etab.push_back(0);
        fp.utab.push_back((unsigned)fp.ukj.size());
      }
      if (fp.ukj.size() > 0xFFFFFFF0ull) {
        fp.gpass.clear(); fp.pmeta.clear(); fp.etab.clear(); fp.utab.clear(); fp.ukj.clear(); fp.ustep.clear();
        return;
      }
      const unsigned info = (unsigned)lanes | (unsigned)steps << 8 | (prev_lev >= 0 && lev != prev_lev ? 1u << 16 : 0u);
      const unsigned meta[8] = {(unsigned)(smask & 0xFFFFFFFFull), (unsigned)(smask >> 32),
                                (unsigned)(dmask & 0xFFFFFFFFull), (unsigned)(dmask >> 32), info, 0u, 0u, 0u};
      fp.pmeta.insert(fp.pmeta.end(), meta, meta + 8);
      prev_lev = lev;
    }
    fp.gpass.push_back((int)(fp.pmeta.size() / 8));
  }
  fp.utab.push_back((unsigned)fp.ukj.size());
  fp.ok = true;
}

void build_sweeps(Sys& s) {
  build_factor_pattern(s);
  if (s.color_of.empty()) {
    level_sweeps(s);
    build_slabs(s, s.fwd, true);
    build_slabs(s, s.bwd, false);
    s.fwd.fuse_bgrp.clear();
    build_factor_plan(s);
    return;
  }
  // contiguous groups: tiles (BMC) or 64-row chunks of each colour (MC)
  std::vector<int> gstart(1, 0);
  const int n = s.n;
  int i = 0;
  while (i < n) {
    const int c = s.color_of[i];
    const int t = s.tile_of.empty() ? -1 : s.tile_of[i];
    int j = i + 1;
    while (j < n && s.color_of[j] == c && (t < 0 ? (j - i) < kGroup : s.tile_of[j] == t)) j++;
    if (j - i > kGroup) throw std::runtime_error("ordering: tile larger than 64 rows");
    gstart.push_back(j);
    i = j;
  }
  group_sweeps(s, gstart);
  build_slabs(s, s.fwd, true);
  build_slabs(s, s.bwd, false);
  fuse_last_first(s);
  build_factor_plan(s);
}

// COrdering_Blk's orderings (src/matvec/ordering_blk.cpp), on the block pattern without the diagonal.
// RCM (MakeOrdering_RCM, :62-134): breadth-first levels from block 0; every level is taken in ascending degree
// (std::multimap keyed by degree keeps equal degrees in the order they were reached), each block numbered as it is
// taken and its unvisited neighbours (ascending) appended to the next level; the numbers are then reversed, and the
// array slot of block b holds b's new number, which the reference then reads as NewToOld.
// RCM2 (MakeOrdering_RCM2, :169-262): level 0 the blocks without lower neighbours; a candidate of the next level
// joins it only when none of its lower neighbours is unnumbered or in the current level (else it leaves the level
// and may be reached again); new -> old in the order blocks are taken.
// AMD (MakeOrdering_AMD, :320-485): repeatedly the uneliminated block of least degree, lowest index first; the
// degree updates there loop over element lists that its commented-out AroundEnode call would fill, so the degrees
// stay the initial ones.
bool make_reference_ordering(int kind, int n, const unsigned* rowptr, const unsigned* colind, std::vector<int>& order) {
  order.assign((size_t)n, -1);
  auto deg = [&](int i) { return rowptr[i + 1] - rowptr[i]; };
  if (kind == DFM_ORDER_NATURAL) {
    std::iota(order.begin(), order.end(), 0);
    return true;
  }
  if (kind == DFM_ORDER_RCM) {
    if (n == 0) return true;
    std::vector<int> level((size_t)n, -1), next, cur;
    level[0] = 0;
    next.push_back(0);
    int taken = 0, lev = 0;
    while (!next.empty()) {
      cur.assign(next.begin(), next.end());
      std::stable_sort(cur.begin(), cur.end(), [&](int a, int b) { return deg(a) < deg(b); });
      next.clear();
      for (int b : cur) {
        order[b] = taken++;
        for (unsigned q = rowptr[b]; q < rowptr[b + 1]; q++) {
          const int j = (int)colind[q];
          if (level[j] != -1) continue;
          level[j] = lev + 1;
          next.push_back(j);
        }
      }
      lev++;
    }
    if (taken != n) return false;  // "isolated island": the reference asserts
    for (int b = 0; b < n; b++) order[b] = n - 1 - order[b];
    return true;
  }
  if (kind == DFM_ORDER_RCM2) {
    std::vector<int> level((size_t)n, -1), cand, next;
    int taken = 0, lev = 0;
    for (int b = 0; b < n; b++) {
      bool lower = false;
      for (unsigned q = rowptr[b]; q < rowptr[b + 1]; q++)
        if ((int)colind[q] < b) { lower = true; break; }
      if (lower) continue;
      level[b] = lev;
      order[taken++] = b;
      for (unsigned q = rowptr[b]; q < rowptr[b + 1]; q++) {
        const int j = (int)colind[q];
        if (level[j] == lev || level[j] == lev + 1) continue;
        cand.push_back(j);
        level[j] = lev + 1;
      }
    }
    lev++;
    while (!cand.empty()) {
      next.clear();
      for (int b : cand) {
        bool blocked = false;
        for (unsigned q = rowptr[b]; q < rowptr[b + 1]; q++) {
          const int j = (int)colind[q];
          if (j < b && (level[j] == lev || level[j] == -1)) { blocked = true; break; }
        }
        if (blocked) { level[b] = -1; continue; }
        next.push_back(b);
      }
      cand.clear();
      for (int b : next) {
        order[taken++] = b;
        for (unsigned q = rowptr[b]; q < rowptr[b + 1]; q++) {
          const int j = (int)colind[q];
          if (level[j] != -1 || j < b) continue;
          level[j] = lev + 1;
          cand.push_back(j);
        }
      }
      lev++;
    }
    return taken == n;
  }
  if (kind == DFM_ORDER_AMD) {
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return deg(a) < deg(b); });
    return true;
  }
  return false;
}

// Greedy node colouring in natural order; order = colour-major, natural inside a colour.
void make_multicolor(const Sys& s, std::vector<int>& order, std::vector<int>& color_of_new, int& ncolors) {
  const int n = s.n;
  std::vector<int> col((size_t)n, -1);
  std::vector<int> used;
  ncolors = 0;
  for (int i = 0; i < n; i++) {
    used.assign((size_t)ncolors + 1, 0);
    for (unsigned q = s.rowptr0[i]; q < s.rowptr0[i + 1]; q++) {
      const int c = col[s.colind0[q]];
      if (c >= 0) used[c] = 1;
    }
    int c = 0;
    while (used[c]) c++;
    col[i] = c;
    ncolors = std::max(ncolors, c + 1);
  }
  order.resize(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return col[a] < col[b]; });
  color_of_new.resize(n);
  for (int k = 0; k < n; k++) color_of_new[k] = col[order[k]];
}

// Block multicolour ordering. Nodes are binned into geometric tiles of about `tedge` element spacings per
// axis (8 x 8 nodes in 2D, 4 x 4 x 4 in 3D by default, so a tile fits one wavefront and its intra-tile
// dependency chain stays short); a tile holding more than 64 nodes is split into chunks of 64 in natural
// order. Tiles are coloured greedily on the tile adjacency graph, in the lexicographic order of their
// (z, y, x) bins; the order is colour-major, tile-major, natural inside a tile. Any such order is a
// permutation that CPreconditioner_ILU::SetOrdering accepts, so the factorisation it induces is the
// reference's ILU(0) under that ordering.
void make_blockcolor(const Sys& s, int tedge, std::vector<int>& order, std::vector<int>& color_of_new,
                     std::vector<int>& tile_of_new, int& ncolors) {
  const int n = s.n, nd = s.ndim, npoel = s.npoel;
  if (tedge <= 0) tedge = (nd == 2) ? 8 : 4;
  // element spacing from the mean element measure (right triangles of h^2/2, Kuhn tets of h^3/6)
  double meas = 0.0;
  for (int e = 0; e < s.nelem; e++) {
    const int* no = &s.lnods[(size_t)e * npoel];
    const double* p0 = &s.coords[(size_t)no[0] * nd];
    double d[3][3];
    for (int k = 1; k < npoel; k++)
      for (int c = 0; c < nd; c++) d[k - 1][c] = s.coords[(size_t)no[k] * nd + c] - p0[c];
    if (nd == 2) meas += 0.5 * std::fabs(d[0][0] * d[1][1] - d[0][1] * d[1][0]);
    else
      meas += std::fabs(d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) - d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
                        d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0])) / 6.0;
  }
  meas /= std::max(1, s.nelem);
  const double h = (nd == 2) ? std::sqrt(2.0 * meas) : std::cbrt(6.0 * meas);
  double lo[3] = {1e300, 1e300, 1e300};
  for (int i = 0; i < n; i++)
    for (int c = 0; c < nd; c++) lo[c] = std::min(lo[c], s.coords[(size_t)i * nd + c]);
  std::vector<long long> key(n);
  for (int i = 0; i < n; i++) {
    long long t[3] = {0, 0, 0};
    for (int c = 0; c < nd; c++) {
      const long long g = (long long)std::floor((s.coords[(size_t)i * nd + c] - lo[c]) / h + 0.5);
      t[c] = g / tedge;
    }
    key[i] = (t[2] << 42) | (t[1] << 21) | t[0];
  }
  // tiles in lexicographic (z, y, x) bin order; nodes in natural order inside a tile, split at 64
  std::vector<int> byk(n);
  std::iota(byk.begin(), byk.end(), 0);
  std::stable_sort(byk.begin(), byk.end(), [&](int a, int b) { return key[a] < key[b]; });
  std::vector<int> tile(n);
  int ntile = 0;
  for (int k = 0; k < n;) {
    int e = k;
    while (e < n && key[byk[e]] == key[byk[k]]) e++;
    for (int c = k; c < e; c += kGroup) {
      for (int j = c; j < std::min(e, c + kGroup); j++) tile[byk[j]] = ntile;
      ntile++;
    }
    k = e;
  }
  std::vector<std::vector<int>> tadj((size_t)ntile);
  for (int i = 0; i < n; i++)
    for (unsigned q = s.rowptr0[i]; q < s.rowptr0[i + 1]; q++) {
      const int a = tile[i], c = tile[s.colind0[q]];
      if (a != c) tadj[a].push_back(c);
    }
  for (auto& a : tadj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  // greedy in tile order, then DSATUR (the uncoloured tile with the most distinct neighbour colours next,
  // ties to the larger degree, then the lower index); the one with fewer colours is kept (fewer sweep phases)
  std::vector<int> tcol((size_t)ntile, -1);
  ncolors = 0;
  std::vector<char> used;
  for (int t = 0; t < ntile; t++) {
    used.assign((size_t)ncolors + 1, 0);
    for (int c : tadj[t])
      if (tcol[c] >= 0) used[tcol[c]] = 1;
    int c = 0;
    while (used[c]) c++;
    tcol[t] = c;
    ncolors = std::max(ncolors, c + 1);
  }
  {
    std::vector<int> dcol((size_t)ntile, -1), sat((size_t)ntile, 0);
    std::vector<std::vector<char>> seen((size_t)ntile);
    std::set<std::tuple<int, int, int>> q;  // (-saturation, -degree, tile)
    for (int t = 0; t < ntile; t++) q.insert(std::make_tuple(0, -(int)tadj[t].size(), t));
    int nc = 0;
    while (!q.empty()) {
      const int t = std::get<2>(*q.begin());
      q.erase(q.begin());
      used.assign((size_t)nc + 1, 0);
      for (int c : tadj[t])
        if (dcol[c] >= 0) used[dcol[c]] = 1;
      int c = 0;
      while (used[c]) c++;
      dcol[t] = c;
      nc = std::max(nc, c + 1);
      for (int u : tadj[t]) {
        if (dcol[u] >= 0) continue;
        auto& sv = seen[u];
        if ((int)sv.size() <= c) sv.resize((size_t)c + 1, 0);
        if (sv[c]) continue;
        sv[c] = 1;
        q.erase(std::make_tuple(-sat[u], -(int)tadj[u].size(), u));
        sat[u]++;
        q.insert(std::make_tuple(-sat[u], -(int)tadj[u].size(), u));
      }
    }
    if (nc < ncolors) {
      tcol.swap(dcol);
      ncolors = nc;
    }
  }
  order = byk;  // already tile-major (tiles in creation order), natural inside a tile
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tcol[tile[a]] < tcol[tile[b]]; });
  color_of_new.resize(n);
  tile_of_new.resize(n);
  for (int k = 0; k < n; k++) {
    color_of_new[k] = tcol[tile[order[k]]];
    tile_of_new[k] = tile[order[k]];
  }
}

}  // namespace dfm
