// Driver that exposes the reference's own host build of TA-Lib 0.8.3
// (src/compat/Talib08Formulas.h + talib08_formulas.inc, compiled for the host
// when __CUDACC__ is not defined) through plain C entry points.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/ from
// the sources where they lie under the reference checkout; it is the
// independent check of oracle/oracle.c and an optional CPU baseline.  Nothing
// here is product code.
//
// Every wrapper widens float32 inputs to double, runs the TA_* function over
// the whole array and returns a full-length float32 line with NaN before
// outBegIdx, which is the reference ABI's layout.
#include "Talib08Formulas.h"

#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

using namespace tacuda::talib08::ref;

namespace {
thread_local std::vector<unsigned char> g_arena_mem;

Arena make_arena(size_t n) {
  size_t need = n * 64 + (1u << 20);
  if (g_arena_mem.size() < need) g_arena_mem.resize(need);
  return Arena{g_arena_mem.data(), g_arena_mem.size()};
}

std::vector<double> widen(const float* x, int n) { return std::vector<double>(x, x + n); }

void place(float* out, int n, int beg, int nb, const double* v) {
  for (int i = 0; i < n; ++i) out[i] = NAN;
  for (int i = 0; i < nb; ++i) out[beg + i] = static_cast<float>(v[i]);
}
}  // namespace

#define API extern "C" __attribute__((visibility("default")))

#define REF_1IN_1OUT(NAME, CALL)                                  \
  API int ref_##NAME(const float* x, float* out, int n, int p) {  \
    std::vector<double> in = widen(x, n), o(n);                   \
    Arena a = make_arena(n);                                      \
    int beg = 0, nb = 0;                                          \
    int rc = CALL;                                                \
    if (rc) return rc;                                            \
    place(out, n, beg, nb, o.data());                             \
    return 0;                                                     \
  }

REF_1IN_1OUT(sma, TA_SMA(a, 0, n - 1, in.data(), p, &beg, &nb, o.data()))
REF_1IN_1OUT(ema, TA_EMA(a, 0, n - 1, in.data(), p, &beg, &nb, o.data()))
REF_1IN_1OUT(wma, TA_WMA(a, 0, n - 1, in.data(), p, &beg, &nb, o.data()))
REF_1IN_1OUT(dema, TA_DEMA(a, 0, n - 1, in.data(), p, &beg, &nb, o.data()))
REF_1IN_1OUT(tema, TA_TEMA(a, 0, n - 1, in.data(), p, &beg, &nb, o.data()))
REF_1IN_1OUT(kama, TA_KAMA(a, 0, n - 1, in.data(), p, &beg, &nb, o.data()))
REF_1IN_1OUT(rsi, TA_RSI(a, 0, n - 1, in.data(), p, &beg, &nb, o.data()))
REF_1IN_1OUT(rma, TA_RMA(a, 0, n - 1, in.data(), p, &beg, &nb, o.data()))
REF_1IN_1OUT(var, TA_VAR(a, 0, n - 1, in.data(), p, 1.0, &beg, &nb, o.data()))
REF_1IN_1OUT(stddev, TA_STDDEV(a, 0, n - 1, in.data(), p, 1.0, &beg, &nb, o.data()))

API int ref_t3(const float* x, float* out, int n, int p, double v) {
  std::vector<double> in = widen(x, n), o(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_T3(a, 0, n - 1, in.data(), p, v, &beg, &nb, o.data());
  if (rc) return rc;
  place(out, n, beg, nb, o.data());
  return 0;
}

API int ref_macd(const float* x, float* m, float* s, float* h, int n, int fast, int slow,
                 int sig) {
  std::vector<double> in = widen(x, n), om(n), os(n), oh(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_MACD(a, 0, n - 1, in.data(), fast, slow, sig, &beg, &nb, om.data(), os.data(),
                   oh.data());
  if (rc) return rc;
  place(m, n, beg, nb, om.data());
  place(s, n, beg, nb, os.data());
  place(h, n, beg, nb, oh.data());
  return 0;
}

API int ref_atr(const float* hi, const float* lo, const float* cl, float* out, int n, int p) {
  std::vector<double> h = widen(hi, n), l = widen(lo, n), c = widen(cl, n), o(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_ATR(a, 0, n - 1, h.data(), l.data(), c.data(), p, &beg, &nb, o.data());
  if (rc) return rc;
  place(out, n, beg, nb, o.data());
  return 0;
}

API int ref_trange(const float* hi, const float* lo, const float* cl, float* out, int n) {
  std::vector<double> h = widen(hi, n), l = widen(lo, n), c = widen(cl, n), o(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_TRANGE(a, 0, n - 1, h.data(), l.data(), c.data(), &beg, &nb, o.data());
  if (rc) return rc;
  place(out, n, beg, nb, o.data());
  return 0;
}

API int ref_bbands(const float* x, float* up, float* mid, float* lo, int n, int p, double ku,
                   double kd) {
  std::vector<double> in = widen(x, n), u(n), m(n), l(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_BBANDS(a, 0, n - 1, in.data(), p, ku, kd, 0, &beg, &nb, u.data(), m.data(),
                     l.data());
  if (rc) return rc;
  place(up, n, beg, nb, u.data());
  place(mid, n, beg, nb, m.data());
  place(lo, n, beg, nb, l.data());
  return 0;
}

API int ref_stoch(const float* hi, const float* lo, const float* cl, float* k, float* d, int n,
                  int kp, int sp, int dp) {
  std::vector<double> h = widen(hi, n), l = widen(lo, n), c = widen(cl, n), ok(n), od(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_STOCH(a, 0, n - 1, h.data(), l.data(), c.data(), kp, sp, 0, dp, 0, &beg, &nb,
                    ok.data(), od.data());
  if (rc) return rc;
  place(k, n, beg, nb, ok.data());
  place(d, n, beg, nb, od.data());
  return 0;
}

// Neighbours that pin oracle functions the vendored library has no exact
// counterpart for: DONCHIAN's bands are MAX(high) and MIN(low); STOCHF's
// fast %K minus 100 is WILLR.
API int ref_donchian(const float* hi, const float* lo, float* up, float* dn, int n, int p) {
  std::vector<double> h = widen(hi, n), l = widen(lo, n), u(n), m(n), d(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_DONCHIAN(a, 0, n - 1, h.data(), l.data(), p, &beg, &nb, u.data(), m.data(),
                       d.data());
  if (rc) return rc;
  place(up, n, beg, nb, u.data());
  place(dn, n, beg, nb, d.data());
  return 0;
}

API int ref_stochf(const float* hi, const float* lo, const float* cl, float* k, float* d, int n,
                   int kp, int dp) {
  std::vector<double> h = widen(hi, n), l = widen(lo, n), c = widen(cl, n), ok(n), od(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_STOCHF(a, 0, n - 1, h.data(), l.data(), c.data(), kp, dp, 0, &beg, &nb, ok.data(),
                     od.data());
  if (rc) return rc;
  place(k, n, beg, nb, ok.data());
  place(d, n, beg, nb, od.data());
  return 0;
}

// TA_MACDEXT with a moving-average type per line (TA_MAType numbering:
// 0 SMA, 1 EMA, 2 WMA, 3 DEMA, 4 TEMA, 5 TRIMA, 6 KAMA, 7 MAMA, 8 T3).
API int ref_macdext(const float* x, float* m, float* s, float* h, int n, int fast, int ft,
                    int slow, int st, int sig, int sgt) {
  std::vector<double> in = widen(x, n), om(n), os(n), oh(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_MACDEXT(a, 0, n - 1, in.data(), fast, ft, slow, st, sig, sgt, &beg, &nb, om.data(),
                      os.data(), oh.data());
  if (rc) return rc;
  place(m, n, beg, nb, om.data());
  place(s, n, beg, nb, os.data());
  place(h, n, beg, nb, oh.data());
  return 0;
}

API int ref_ma(const float* x, float* out, int n, int p, int type) {
  std::vector<double> in = widen(x, n), o(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_MA(a, 0, n - 1, in.data(), p, type, &beg, &nb, o.data());
  if (rc) return rc;
  place(out, n, beg, nb, o.data());
  return 0;
}

// Price transforms, ROC / ROCP, MIDPRICE and LINEARREG of the vendored
// TA-Lib 0.8.3 (checks of the oracle's restatements).
REF_1IN_1OUT(roc, TA_ROC(a, 0, n - 1, in.data(), p, &beg, &nb, o.data()))
REF_1IN_1OUT(rocp, TA_ROCP(a, 0, n - 1, in.data(), p, &beg, &nb, o.data()))
REF_1IN_1OUT(linearreg, TA_LINEARREG(a, 0, n - 1, in.data(), p, &beg, &nb, o.data()))

API int ref_medprice(const float* hi, const float* lo, float* out, int n) {
  std::vector<double> h = widen(hi, n), l = widen(lo, n), o(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_MEDPRICE(a, 0, n - 1, h.data(), l.data(), &beg, &nb, o.data());
  if (rc) return rc;
  place(out, n, beg, nb, o.data());
  return 0;
}

API int ref_typprice(const float* hi, const float* lo, const float* cl, float* out, int n) {
  std::vector<double> h = widen(hi, n), l = widen(lo, n), c = widen(cl, n), o(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_TYPPRICE(a, 0, n - 1, h.data(), l.data(), c.data(), &beg, &nb, o.data());
  if (rc) return rc;
  place(out, n, beg, nb, o.data());
  return 0;
}

API int ref_midprice(const float* hi, const float* lo, float* out, int n, int p) {
  std::vector<double> h = widen(hi, n), l = widen(lo, n), o(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_MIDPRICE(a, 0, n - 1, h.data(), l.data(), p, &beg, &nb, o.data());
  if (rc) return rc;
  place(out, n, beg, nb, o.data());
  return 0;
}

API int ref_ibs(const float* hi, const float* lo, const float* cl, float* out, int n) {
  std::vector<double> h = widen(hi, n), l = widen(lo, n), c = widen(cl, n), o(n);
  Arena a = make_arena(n);
  int beg = 0, nb = 0;
  int rc = TA_IBS(a, 0, n - 1, h.data(), l.data(), c.data(), &beg, &nb, o.data());
  if (rc) return rc;
  place(out, n, beg, nb, o.data());
  return 0;
}

// The reference's 0.8.3 backend around one TA_* call (src/compat/Talib08.cu
// :72-99): skip the leading bars where any input is NaN (keeping the last
// bar), run the routine from index 0 over the rest, place its outputs at
// prefix + outBegIdx.  name selects the routine; p0/p1 are its periods.
API int ref_rows(const char* name, const float* const* in, float* const* out, int n, int p0,
                 int p1, double d0) {
  struct Spec {
    const char* name;
    int nin, nout;
  };
  static const Spec specs[] = {{"cumsum", 1, 1}, {"pvt", 2, 1},      {"rma", 1, 1},
                               {"zlema", 1, 1},  {"efi", 2, 1},      {"vwma", 2, 1},
                               {"er", 1, 1},     {"vhf", 1, 1},      {"dpo", 1, 1},
                               {"chop", 3, 1},   {"adr", 2, 1},      {"ao", 2, 1},
                               {"qstick", 2, 1}, {"donchian", 2, 3}, {"marketfi", 3, 1},
                               {"wad", 3, 1},    {"mcgd", 1, 1},     {"rvol", 1, 1},
                               {"cmf", 4, 1},    {"emv", 3, 1},      {"fosc", 1, 1},
                               {"eri", 3, 2},    {"percentrank", 1, 1}, {"si", 4, 1},
                               {"asi_wilder", 4, 1}, {"vortex", 3, 2}, {"cmou", 1, 1}};
  const Spec* s = nullptr;
  for (const Spec& q : specs)
    if (std::string(q.name) == name) s = &q;
  if (!s || n <= 0) return 4;
  for (int k = 0; k < s->nout; ++k)
    for (int t = 0; t < n; ++t) out[k][t] = NAN;
  int f = 0;
  for (; f < n - 1; ++f) {
    bool any = false;
    for (int j = 0; j < s->nin; ++j) any |= std::isnan(in[j][f]);
    if (!any) break;
  }
  const int m = n - f;
  std::vector<std::vector<double>> x(s->nin), y(s->nout, std::vector<double>(m));
  for (int j = 0; j < s->nin; ++j) x[j].assign(in[j] + f, in[j] + n);
  Arena a = make_arena(m);
  int beg = 0, nb = 0, rc = 0;
  const std::string nm(name);
  const int e = m - 1;
  if (nm == "cumsum") rc = TA_CUMSUM(a, 0, e, x[0].data(), &beg, &nb, y[0].data());
  else if (nm == "pvt") rc = TA_PVT(a, 0, e, x[0].data(), x[1].data(), &beg, &nb, y[0].data());
  else if (nm == "rma") rc = TA_RMA(a, 0, e, x[0].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "zlema") rc = TA_ZLEMA(a, 0, e, x[0].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "efi") rc = TA_EFI(a, 0, e, x[0].data(), x[1].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "vwma") rc = TA_VWMA(a, 0, e, x[0].data(), x[1].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "er") rc = TA_ER(a, 0, e, x[0].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "vhf") rc = TA_VHF(a, 0, e, x[0].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "dpo") rc = TA_DPO(a, 0, e, x[0].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "chop")
    rc = TA_CHOP(a, 0, e, x[0].data(), x[1].data(), x[2].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "adr") rc = TA_ADR(a, 0, e, x[0].data(), x[1].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "ao")
    rc = TA_AO(a, 0, e, x[0].data(), x[1].data(), p0, p1, &beg, &nb, y[0].data());
  else if (nm == "qstick")
    rc = TA_QSTICK(a, 0, e, x[0].data(), x[1].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "donchian")
    rc = TA_DONCHIAN(a, 0, e, x[0].data(), x[1].data(), p0, &beg, &nb, y[0].data(), y[1].data(),
                     y[2].data());
  else if (nm == "marketfi")
    rc = TA_MARKETFI(a, 0, e, x[0].data(), x[1].data(), x[2].data(), &beg, &nb, y[0].data());
  else if (nm == "wad")
    rc = TA_WAD(a, 0, e, x[0].data(), x[1].data(), x[2].data(), &beg, &nb, y[0].data());
  else if (nm == "mcgd") rc = TA_MCGD(a, 0, e, x[0].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "rvol") rc = TA_RVOL(a, 0, e, x[0].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "cmf")
    rc = TA_CMF(a, 0, e, x[0].data(), x[1].data(), x[2].data(), x[3].data(), p0, &beg, &nb,
                y[0].data());
  else if (nm == "emv")
    rc = TA_EMV(a, 0, e, x[0].data(), x[1].data(), x[2].data(), p0, d0, &beg, &nb, y[0].data());
  else if (nm == "fosc") rc = TA_FOSC(a, 0, e, x[0].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "eri")
    rc = TA_ERI(a, 0, e, x[0].data(), x[1].data(), x[2].data(), p0, &beg, &nb, y[0].data(),
                y[1].data());
  else if (nm == "percentrank")
    rc = TA_PERCENTRANK(a, 0, e, x[0].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "si")
    rc = TA_SI(a, 0, e, x[0].data(), x[1].data(), x[2].data(), x[3].data(), d0, &beg, &nb,
               y[0].data());
  else if (nm == "vortex")
    rc = TA_VORTEX(a, 0, e, x[0].data(), x[1].data(), x[2].data(), p0, &beg, &nb, y[0].data(),
                   y[1].data());
  else if (nm == "cmou") rc = TA_CMOU(a, 0, e, x[0].data(), p0, &beg, &nb, y[0].data());
  else if (nm == "asi_wilder")
    rc = TA_ASI(a, 0, e, x[0].data(), x[1].data(), x[2].data(), x[3].data(), d0, &beg, &nb,
                y[0].data());
  if (rc) return 4;
  for (int k = 0; k < s->nout; ++k)
    for (int i = 0; i < nb; ++i) out[k][f + beg + i] = (float)y[k][i];
  return 0;
}
