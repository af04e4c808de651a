// hj_strat.cpp -- strategy generators (own constructions of the reference's
// families) and the strat.so loader with the reference's semantics.
#include "hj_strat.hpp"
#include "hj_util.hpp"

#include <dlfcn.h>
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/hypjac.h"

extern "C" {
unsigned STRAT0 = 0u, STRAT0_STEPS = 0u, STRAT0_PAIRS = 0u;
unsigned STRAT1 = 0u, STRAT1_STEPS = 0u, STRAT1_PAIRS = 0u;
unsigned char strat0[STRAT0_MAX_STEPS][STRAT0_MAX_PAIRS][2];
unsigned short strat1[STRAT1_MAX_STEPS][STRAT1_MAX_PAIRS][2];
}

namespace hj {

const char *const kFamilies[6] = {"mmstep", "BrentL", "colcyc", "cycloc", "rowcyc", "cycwor"};

unsigned family_id(const char *fam)
{
  for (unsigned i = 0u; i < 6u; ++i)
    if (!strcmp(fam, kFamilies[i]))
      return i + 1u;
  return 0u;
}

unsigned family_steps(const char *fam, unsigned n)
{
  const unsigned id = family_id(fam);
  if (!id)
    return 0u;
  return (id == STRAT_MMSTEP) ? n : (n - 1u);
}

static bool pow2(unsigned n) { return n >= 2u && !(n & (n - 1u)); }

bool can_generate(const char *fam, unsigned n)
{
  const unsigned id = family_id(fam);
  if (!id || n < 2u || n > 1024u || (n & 1u))
    return false;
  if (id == STRAT_MMSTEP || id == STRAT_BRENTL)
    return true;
  return pow2(n);  // cycwor, rowcyc, colcyc, cycloc
}

namespace {

using Table = std::vector<std::vector<std::pair<unsigned, unsigned>>>;

// Cyclic strategy without reversal ("cycwor") for n = 2^k: first the n/2
// steps that pair every column of the lower half with every column of the
// upper half (column i meets n-1-(i xor s) in step s), then the two halves
// run the order-n/2 strategy side by side.
Table gen_cycwor(unsigned n)
{
  Table t;
  if (n == 2u) {
    t.push_back({{0u, 1u}});
    return t;
  }
  const unsigned h = n / 2u;
  for (unsigned s = 0u; s < h; ++s) {
    std::vector<std::pair<unsigned, unsigned>> st;
    for (unsigned i = 0u; i < h; ++i)
      st.push_back({i, n - 1u - (i ^ s)});
    t.push_back(st);
  }
  const Table sub = gen_cycwor(h);
  for (const auto &ss : sub) {
    std::vector<std::pair<unsigned, unsigned>> st(ss);
    for (const auto &pq : ss)
      st.push_back({pq.first + h, pq.second + h});
    t.push_back(st);
  }
  return t;
}

// Order doubling of a cyclic strategy (the construction of
// strat/n2n/n2n.sh.cpp:39-128): step 0 pairs (2j, 2j+1); each step of the
// order-n/2 table with pair (i, j) in slot a becomes two steps with slots
// 2a, 2a+1: first (2i,2j),(2i+1,2j+1), then the cross pairs, whose order in
// the two slots distinguishes the row-cyclic from the column-cyclic family.
Table gen_double(const Table &half, bool colcyc)
{
  const unsigned np = (unsigned)half[0].size();  // pairs of the half table
  const unsigned N = 2u * np;                     // order of the half table
  Table t;
  std::vector<std::pair<unsigned, unsigned>> s0;
  for (unsigned j = 0u; j < N; ++j)
    s0.push_back({2u * j, 2u * j + 1u});
  t.push_back(s0);
  for (const auto &hs : half) {
    std::vector<std::pair<unsigned, unsigned>> a, b;
    for (const auto &pq : hs) {
      const unsigned i = pq.first, j = pq.second;
      a.push_back({2u * i, 2u * j});
      a.push_back({2u * i + 1u, 2u * j + 1u});
      if (colcyc) {
        b.push_back({2u * i + 1u, 2u * j});
        b.push_back({2u * i, 2u * j + 1u});
      }
      else {
        b.push_back({2u * i, 2u * j + 1u});
        b.push_back({2u * i + 1u, 2u * j});
      }
    }
    t.push_back(a);
    t.push_back(b);
  }
  return t;
}

Table gen_rowcol(unsigned n, bool colcyc)
{
  Table t = {{{0u, 1u}}};
  for (unsigned m = 4u; m <= n; m *= 2u)
    t = gen_double(t, colcyc);
  return t;
}

// Brent-Luk caterpillar ordering (strat/BrentL.f90:43-103).  Slot k holds
// the column pair (L_k, R_k); after each step the right column of slot 1
// moves to the left of slot 2, each other left column moves one slot right,
// each right column one slot left, the last slot's left column turns into
// its right column, and slot 1 keeps its left column.
Table gen_brentl(unsigned n)
{
  const unsigned h = n / 2u;
  std::vector<unsigned> L(h), R(h);
  for (unsigned k = 0u; k < h; ++k) {
    L[k] = 2u * k;
    R[k] = 2u * k + 1u;
  }
  Table t;
  for (unsigned s = 0u; s + 1u < n; ++s) {
    std::vector<std::pair<unsigned, unsigned>> st;
    for (unsigned k = 0u; k < h; ++k)
      st.push_back(L[k] < R[k] ? std::make_pair(L[k], R[k]) : std::make_pair(R[k], L[k]));
    t.push_back(st);
    if (h == 1u)
      continue;
    std::vector<unsigned> nL(L), nR(R);
    nL[1] = R[0];
    for (unsigned k = 2u; k < h; ++k)
      nL[k] = L[k - 1u];
    for (unsigned k = 0u; k + 1u < h; ++k)
      nR[k] = R[k + 1u];
    nR[h - 1u] = L[h - 1u];
    L.swap(nL);
    R.swap(nR);
  }
  return t;
}

// Modulus strategy: step s pairs the columns whose indices sum to n-1+s
// (mod n); when that sum is even the two columns that would pair with
// themselves, (s-1)/2 and (s-1)/2 + n/2, are paired together.  Pairs are
// listed from (s, n-1) upward, then the self pair, then downward from below s.
Table gen_mmstep(unsigned n)
{
  Table t;
  for (unsigned s = 0u; s < n; ++s) {
    std::vector<std::pair<unsigned, unsigned>> st;
    for (unsigned i = s; (n - 1u + s - i) > i; ++i)
      st.push_back({i, n - 1u + s - i});
    int start;
    if (s & 1u) {
      const unsigned a = (s - 1u) / 2u;
      st.push_back({a, a + n / 2u});
      start = (int)a - 1;
    }
    else
      start = (int)(s / 2u) - 1;
    for (int i = start; i >= 0; --i) {
      const int j = (int)s - 1 - i;
      if (i < j)
        st.push_back({(unsigned)i, (unsigned)j});
    }
    t.push_back(st);
  }
  return t;
}

}  // namespace

bool generate(const char *fam, unsigned n, Strategy &out)
{
  if (!can_generate(fam, n))
    return false;
  Table t;
  switch (family_id(fam)) {
  case STRAT_MMSTEP: t = gen_mmstep(n); break;
  case STRAT_BRENTL: t = gen_brentl(n); break;
  case STRAT_COLCYC: t = gen_rowcol(n, true); break;
  case STRAT_CYCLOC: {
    Table c = gen_rowcol(n, true);
    t.assign(c.rbegin(), c.rend());
    break;
  }
  case STRAT_ROWCYC: t = gen_rowcol(n, false); break;
  case STRAT_CYCWOR: t = gen_cycwor(n); break;
  default: return false;
  }
  out.n = n;
  out.steps = (unsigned)t.size();
  out.pairs = n / 2u;
  out.tab.assign((size_t)out.steps * out.pairs * 2u, 0u);
  for (unsigned s = 0u; s < out.steps; ++s) {
    if (t[s].size() != out.pairs)
      return false;
    for (unsigned p = 0u; p < out.pairs; ++p) {
      out.tab[((size_t)s * out.pairs + p) * 2u + 0u] = (uint16_t)t[s][p].first;
      out.tab[((size_t)s * out.pairs + p) * 2u + 1u] = (uint16_t)t[s][p].second;
    }
  }
  return out.steps == family_steps(fam, n);
}

}  // namespace hj

// ---------------------------------------------------------------------------
// Reference-compatible plugin loader (share/my_utils.cu:14-61,
// share/HypJacL.cu:11-121).

extern "C" int fexist(const char *const fn)
{
  struct stat buf;
  return (fn ? (0 == stat(fn, &buf)) : 0);
}

extern "C" void *strat_open(const char *const sdy)
{
  return dlopen(sdy, RTLD_LAZY);
}

extern "C" int strat_close(void *const h)
{
  return dlclose(h);
}

extern "C" const void *strat_ptr(void *const h, const char *const snp, const unsigned n)
{
  char arrn[12] = {'\0'};
  (void)snprintf(arrn, sizeof(arrn), "%s%05u", snp, n);
  return (arrn[11] ? nullptr : dlsym(h, arrn));
}

extern "C" void init_strats(const char *const sdy, const char *const snp0, const unsigned n0,
                            const char *const snp1, const unsigned n1)
{
  if (!sdy || !*sdy) HJ_DIE("SDY empty");
  if (!snp0 || !*snp0) HJ_DIE("SNP0 empty");
  if (!n0 || (n0 & 1u)) HJ_DIE("n0 not positive even");
  if (!snp1 || !*snp1) HJ_DIE("SNP1 empty");
  if (!n1 || (n1 & 1u)) HJ_DIE("n1 not positive even");

  const unsigned id0 = hj::family_id(snp0);
  if (!id0) HJ_DIE("SNP0 unknown");
  const unsigned id1 = hj::family_id(snp1);
  if (!id1) HJ_DIE("SNP1 unknown");
  if (n0 > 2u * STRAT0_MAX_PAIRS) HJ_DIE("n0 too large");
  if (n1 > 2u * STRAT1_MAX_PAIRS) HJ_DIE("n1 too large");

  STRAT0 = id0;
  STRAT0_STEPS = hj::family_steps(snp0, n0);
  STRAT0_PAIRS = n0 >> 1u;
  (void)memset(strat0, 0, sizeof(strat0));
  STRAT1 = id1;
  STRAT1_STEPS = hj::family_steps(snp1, n1);
  STRAT1_PAIRS = n1 >> 1u;
  (void)memset(strat1, 0, sizeof(strat1));

  void *const h = strat_open(sdy);
  if (!h) {
    (void)fprintf(stderr, "[ERROR] %s(%d): %s\n", __FILE__, __LINE__, dlerror());
    exit(EXIT_FAILURE);
  }
  const unsigned char *a0 = static_cast<const unsigned char *>(strat_ptr(h, snp0, n0));
  if (!a0) {
    (void)strat_close(h);
    HJ_DIE("NO SUCH STRATEGY for SNP0 & n0");
  }
  for (unsigned s = 0u; s < STRAT0_STEPS; ++s)
    for (unsigned p = 0u; p < STRAT0_PAIRS; ++p)
      for (unsigned i = 0u; i < 2u; ++i)
        strat0[s][p][i] = *a0++;
  if (n1 <= 256u) {
    const unsigned char *a1 = static_cast<const unsigned char *>(strat_ptr(h, snp1, n1));
    if (!a1) {
      (void)strat_close(h);
      HJ_DIE("NO SUCH STRATEGY for SNP1 & n1");
    }
    for (unsigned s = 0u; s < STRAT1_STEPS; ++s)
      for (unsigned p = 0u; p < STRAT1_PAIRS; ++p)
        for (unsigned i = 0u; i < 2u; ++i)
          strat1[s][p][i] = *a1++;
  }
  else {
    const unsigned short *a1 = static_cast<const unsigned short *>(strat_ptr(h, snp1, n1));
    if (!a1) {
      (void)strat_close(h);
      HJ_DIE("NO SUCH STRATEGY for SNP1 & n1");
    }
    for (unsigned s = 0u; s < STRAT1_STEPS; ++s)
      for (unsigned p = 0u; p < STRAT1_PAIRS; ++p)
        for (unsigned i = 0u; i < 2u; ++i)
          strat1[s][p][i] = *a1++;
  }
  if (strat_close(h)) HJ_DIE("strat_close failed");
}
