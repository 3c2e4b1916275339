/* CPU restatement of the 61 TA-Lib candlestick patterns (test
 * infrastructure only: the oracle the GPU kernels are checked against).
 *
 * Follows the reference's literal TA-Lib ports,
 * include/utils/CandleUtils.h:1434-2339 (each the decision block of the
 * corresponding ta_CDL*.c), evaluated in double over the float32 inputs,
 * with TA-Lib's own double candle-setting factors (the reference keeps them
 * in float, CandleUtils.h:37-64; TA-Lib's ta_global.c defaults are doubles).
 * Averages are TA_CANDLEAVERAGE: factor * (sum of the setting's range over
 * the avgPeriod candles strictly before the candle) / avgPeriod, halved for
 * the Shadows range; settings with avgPeriod 0 use the candle itself.
 * Output: 0 for bars before the pattern's lookback (reference
 * src/indicators/<Pattern>.cu warm-ups), the signed TA-Lib signal after.
 *
 * TA-Lib itself is not importable here.  Every pattern is pinned by value:
 * 10 by the reference's hand-authored activation fixture
 * (differential_corpus.py:73-142), 6 by independent numpy restatements
 * (tests/test_oracle_candles.py) and the other 45 by constructive fixtures
 * built from TA-Lib's rule statements, each a sequence that fires with a
 * stated signal plus near misses that break one condition by one tick
 * (tests/test_oracle_candle_fixtures.py); signal-set and symmetry
 * invariants on random bars besides. */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define API __attribute__((visibility("default")))

enum { RB_ = 0, HL_ = 1, SH_ = 2 };
typedef struct {
  int range, period;
  double factor;
} Setting;
enum { BodyLong, BodyVeryLong, BodyShort, BodyDoji, ShadowLong, ShadowVeryLong, ShadowShort,
       ShadowVeryShort, Near, Far, Equal };
static const Setting SET[] = {
    {RB_, 10, 1.0}, {RB_, 10, 3.0}, {RB_, 10, 1.0}, {HL_, 10, 0.1}, {RB_, 0, 1.0},
    {RB_, 0, 2.0},  {SH_, 10, 1.0}, {HL_, 10, 0.1}, {HL_, 5, 0.2},  {HL_, 5, 0.6},
    {HL_, 5, 0.05}};

typedef struct {
  const float *o, *h, *l, *c;
} Bars;

static double range_of(const Bars* b, int kind, int j) {
  const double o = b->o[j], h = b->h[j], l = b->l[j], c = b->c[j];
  if (kind == RB_) return fabs(c - o);
  if (kind == HL_) return h - l;
  return (h - fmax(o, c)) + (fmin(o, c) - l);
}

static double avg(const Bars* b, int s, int j) {
  const Setting* st = &SET[s];
  const double div = st->range == SH_ ? 2.0 : 1.0;
  if (st->period == 0) return st->factor * range_of(b, st->range, j) / div;
  double sum = 0.0;
  for (int u = j - st->period; u < j; ++u) sum += range_of(b, st->range, u);
  return st->factor * (sum / st->period) / div;
}

#define O(j) ((double)B->o[(j)])
#define H(j) ((double)B->h[(j)])
#define L(j) ((double)B->l[(j)])
#define C(j) ((double)B->c[(j)])
#define RB(j) fabs(C(j) - O(j))
#define US(j) (H(j) - fmax(O(j), C(j)))
#define LS(j) (fmin(O(j), C(j)) - L(j))
#define COL(j) (C(j) >= O(j) ? 1 : -1)
#define AVG(S, j) avg(B, S, (j))
#define TOP(j) fmax(O(j), C(j))
#define BOT(j) fmin(O(j), C(j))
#define RBGAPUP(a, p) (BOT(a) > TOP(p))
#define RBGAPDOWN(a, p) (TOP(a) < BOT(p))
#define GAPUP(a, p) (L(a) > H(p))
#define GAPDOWN(a, p) (H(a) < L(p))
#define P(name) static int name(const Bars* B, int i)

P(belthold) {
  if (RB(i) > AVG(BodyLong, i) && ((COL(i) == 1 && LS(i) < AVG(ShadowVeryShort, i)) ||
                                   (COL(i) == -1 && US(i) < AVG(ShadowVeryShort, i))))
    return COL(i) * 100;
  return 0;
}
P(closingmarubozu) {
  if (RB(i) > AVG(BodyLong, i) && ((COL(i) == 1 && US(i) < AVG(ShadowVeryShort, i)) ||
                                   (COL(i) == -1 && LS(i) < AVG(ShadowVeryShort, i))))
    return COL(i) * 100;
  return 0;
}
P(marubozu) {
  if (RB(i) > AVG(BodyLong, i) && US(i) < AVG(ShadowVeryShort, i) &&
      LS(i) < AVG(ShadowVeryShort, i))
    return COL(i) * 100;
  return 0;
}
P(highwave) {
  if (RB(i) < AVG(BodyShort, i) && US(i) > AVG(ShadowVeryLong, i) &&
      LS(i) > AVG(ShadowVeryLong, i))
    return COL(i) * 100;
  return 0;
}
P(longleggeddoji) {
  return (RB(i) <= AVG(BodyDoji, i) &&
          (LS(i) > AVG(ShadowLong, i) || US(i) > AVG(ShadowLong, i))) ? 100 : 0;
}
P(longline) {
  if (RB(i) > AVG(BodyLong, i) && US(i) < AVG(ShadowShort, i) && LS(i) < AVG(ShadowShort, i))
    return COL(i) * 100;
  return 0;
}
P(shortline) {
  if (RB(i) < AVG(BodyShort, i) && US(i) < AVG(ShadowShort, i) && LS(i) < AVG(ShadowShort, i))
    return COL(i) * 100;
  return 0;
}
P(spinningtop) {
  if (RB(i) < AVG(BodyShort, i) && US(i) > RB(i) && LS(i) > RB(i)) return COL(i) * 100;
  return 0;
}
P(rickshawman) {
  const double mid = L(i) + (H(i) - L(i)) / 2.0;
  return (RB(i) <= AVG(BodyDoji, i) && LS(i) > AVG(ShadowLong, i) &&
          US(i) > AVG(ShadowLong, i) && BOT(i) <= mid + AVG(Near, i) &&
          TOP(i) >= mid - AVG(Near, i)) ? 100 : 0;
}
P(takuri) {
  return (RB(i) <= AVG(BodyDoji, i) && US(i) < AVG(ShadowVeryShort, i) &&
          LS(i) > AVG(ShadowVeryLong, i)) ? 100 : 0;
}
P(doji) { return RB(i) <= AVG(BodyDoji, i) ? 100 : 0; }
P(dragonflydoji) {
  return (RB(i) <= AVG(BodyDoji, i) && US(i) < AVG(ShadowVeryShort, i) &&
          LS(i) > AVG(ShadowVeryShort, i)) ? 100 : 0;
}
P(gravestonedoji) {
  return (RB(i) <= AVG(BodyDoji, i) && LS(i) < AVG(ShadowVeryShort, i) &&
          US(i) > AVG(ShadowVeryShort, i)) ? 100 : 0;
}
P(hammer) {
  return (RB(i) < AVG(BodyShort, i) && LS(i) > AVG(ShadowLong, i) &&
          US(i) < AVG(ShadowVeryShort, i) && BOT(i) <= L(i - 1) + AVG(Near, i - 1)) ? 100 : 0;
}
P(invertedhammer) {
  return (RB(i) < AVG(BodyShort, i) && US(i) > AVG(ShadowLong, i) &&
          LS(i) < AVG(ShadowVeryShort, i) && RBGAPDOWN(i, i - 1)) ? 100 : 0;
}
P(hangingman) {
  return (RB(i) < AVG(BodyShort, i) && LS(i) > AVG(ShadowLong, i) &&
          US(i) < AVG(ShadowVeryShort, i) && BOT(i) >= H(i - 1) - AVG(Near, i - 1)) ? -100 : 0;
}
P(shootingstar) {
  return (RB(i) < AVG(BodyShort, i) && US(i) > AVG(ShadowLong, i) &&
          LS(i) < AVG(ShadowVeryShort, i) && RBGAPUP(i, i - 1)) ? -100 : 0;
}
P(homingpigeon) {
  return (COL(i - 1) == -1 && COL(i) == -1 && RB(i - 1) > AVG(BodyLong, i - 1) &&
          RB(i) <= AVG(BodyShort, i) && O(i) < O(i - 1) && C(i) > C(i - 1)) ? 100 : 0;
}
P(matchinglow) {
  return (COL(i - 1) == -1 && COL(i) == -1 && C(i) <= C(i - 1) + AVG(Equal, i - 1) &&
          C(i) >= C(i - 1) - AVG(Equal, i - 1)) ? 100 : 0;
}
P(onneck) {
  return (COL(i - 1) == -1 && RB(i - 1) > AVG(BodyLong, i - 1) && COL(i) == 1 &&
          O(i) < L(i - 1) && C(i) <= L(i - 1) + AVG(Equal, i - 1) &&
          C(i) >= L(i - 1) - AVG(Equal, i - 1)) ? -100 : 0;
}
P(inneck) {
  return (COL(i - 1) == -1 && RB(i - 1) > AVG(BodyLong, i - 1) && COL(i) == 1 &&
          O(i) < L(i - 1) && C(i) <= C(i - 1) + AVG(Equal, i - 1) && C(i) >= C(i - 1)) ? -100 : 0;
}
P(thrusting) {
  return (COL(i - 1) == -1 && RB(i - 1) > AVG(BodyLong, i - 1) && COL(i) == 1 &&
          O(i) < L(i - 1) && C(i) > C(i - 1) + AVG(Equal, i - 1) &&
          C(i) <= C(i - 1) + RB(i - 1) * 0.5) ? -100 : 0;
}
P(separatinglines) {
  if (COL(i - 1) == -COL(i) && O(i) <= O(i - 1) + AVG(Equal, i - 1) &&
      O(i) >= O(i - 1) - AVG(Equal, i - 1) && RB(i) > AVG(BodyLong, i) &&
      ((COL(i) == 1 && LS(i) < AVG(ShadowVeryShort, i)) ||
       (COL(i) == -1 && US(i) < AVG(ShadowVeryShort, i))))
    return COL(i) * 100;
  return 0;
}
P(counterattack) {
  if (COL(i - 1) == -COL(i) && RB(i - 1) > AVG(BodyLong, i - 1) && RB(i) > AVG(BodyLong, i) &&
      C(i) <= C(i - 1) + AVG(Equal, i - 1) && C(i) >= C(i - 1) - AVG(Equal, i - 1))
    return COL(i) * 100;
  return 0;
}
P(gapsidesidewhite) {
  if (((RBGAPUP(i - 1, i - 2) && RBGAPUP(i, i - 2)) ||
       (RBGAPDOWN(i - 1, i - 2) && RBGAPDOWN(i, i - 2))) &&
      COL(i - 1) == 1 && COL(i) == 1 && RB(i) >= RB(i - 1) - AVG(Near, i - 1) &&
      RB(i) <= RB(i - 1) + AVG(Near, i - 1) && O(i) >= O(i - 1) - AVG(Equal, i - 1) &&
      O(i) <= O(i - 1) + AVG(Equal, i - 1))
    return RBGAPUP(i - 1, i - 2) ? 100 : -100;
  return 0;
}
P(upsidegap2crows) {
  return (COL(i - 2) == 1 && RB(i - 2) > AVG(BodyLong, i - 2) && COL(i - 1) == -1 &&
          RB(i - 1) <= AVG(BodyShort, i - 1) && RBGAPUP(i - 1, i - 2) && COL(i) == -1 &&
          O(i) > O(i - 1) && C(i) < C(i - 1) && C(i) > C(i - 2)) ? -100 : 0;
}
P(threeinside) {
  if (RB(i - 2) > AVG(BodyLong, i - 2) && RB(i - 1) <= AVG(BodyShort, i - 1) &&
      TOP(i - 1) < TOP(i - 2) && BOT(i - 1) > BOT(i - 2) &&
      ((COL(i - 2) == 1 && COL(i) == -1 && C(i) < O(i - 2)) ||
       (COL(i - 2) == -1 && COL(i) == 1 && C(i) > O(i - 2))))
    return -COL(i - 2) * 100;
  return 0;
}
P(threeblackcrows) {
  return (COL(i - 3) == 1 && COL(i - 2) == -1 && LS(i - 2) < AVG(ShadowVeryShort, i - 2) &&
          COL(i - 1) == -1 && LS(i - 1) < AVG(ShadowVeryShort, i - 1) && COL(i) == -1 &&
          LS(i) < AVG(ShadowVeryShort, i) && O(i - 1) < O(i - 2) && O(i - 1) > C(i - 2) &&
          O(i) < O(i - 1) && O(i) > C(i - 1) && H(i - 3) > C(i - 2) && C(i - 2) > C(i - 1) &&
          C(i - 1) > C(i)) ? -100 : 0;
}
P(threestarsinsouth) {
  return (COL(i - 2) == -1 && COL(i - 1) == -1 && COL(i) == -1 &&
          RB(i - 2) > AVG(BodyLong, i - 2) && LS(i - 2) > AVG(ShadowLong, i - 2) &&
          RB(i - 1) < RB(i - 2) && O(i - 1) > C(i - 2) && O(i - 1) <= H(i - 2) &&
          L(i - 1) < C(i - 2) && L(i - 1) >= L(i - 2) &&
          LS(i - 1) > AVG(ShadowVeryShort, i - 1) && RB(i) < AVG(BodyShort, i) &&
          LS(i) < AVG(ShadowVeryShort, i) && US(i) < AVG(ShadowVeryShort, i) &&
          L(i) > L(i - 1) && H(i) < H(i - 1)) ? 100 : 0;
}
P(tristar) {
  /* all three doji tests use the average anchored at i-2 (TA-Lib) */
  if (RB(i - 2) <= AVG(BodyDoji, i - 2) && RB(i - 1) <= AVG(BodyDoji, i - 2) &&
      RB(i) <= AVG(BodyDoji, i - 2)) {
    int out = 0;
    if (RBGAPUP(i - 1, i - 2) && TOP(i) < TOP(i - 1)) out = -100;
    if (RBGAPDOWN(i - 1, i - 2) && BOT(i) > BOT(i - 1)) out = 100;
    return out;
  }
  return 0;
}
P(concealbabyswall) {
  return (COL(i - 3) == -1 && COL(i - 2) == -1 && COL(i - 1) == -1 && COL(i) == -1 &&
          LS(i - 3) < AVG(ShadowVeryShort, i - 3) && US(i - 3) < AVG(ShadowVeryShort, i - 3) &&
          LS(i - 2) < AVG(ShadowVeryShort, i - 2) && US(i - 2) < AVG(ShadowVeryShort, i - 2) &&
          RBGAPDOWN(i - 1, i - 2) && US(i - 1) > AVG(ShadowVeryShort, i - 1) &&
          H(i - 1) > C(i - 2) && H(i) > H(i - 1) && L(i) < L(i - 1)) ? 100 : 0;
}
P(abandonedbaby) {
  const double pen = 0.3;
  if (RB(i - 2) > AVG(BodyLong, i - 2) && RB(i - 1) <= AVG(BodyDoji, i - 1) &&
      RB(i) > AVG(BodyShort, i) &&
      ((COL(i - 2) == 1 && COL(i) == -1 && C(i) < C(i - 2) - RB(i - 2) * pen &&
        GAPUP(i - 1, i - 2) && GAPDOWN(i, i - 1)) ||
       (COL(i - 2) == -1 && COL(i) == 1 && C(i) > C(i - 2) + RB(i - 2) * pen &&
        GAPDOWN(i - 1, i - 2) && GAPUP(i, i - 1))))
    return COL(i) * 100;
  return 0;
}
P(advanceblock) {
  return (COL(i - 2) == 1 && COL(i - 1) == 1 && COL(i) == 1 && C(i) > C(i - 1) &&
          C(i - 1) > C(i - 2) && O(i - 1) > O(i - 2) && O(i - 1) <= C(i - 2) + AVG(Near, i - 2) &&
          O(i) > O(i - 1) && O(i) <= C(i - 1) + AVG(Near, i - 1) &&
          RB(i - 2) > AVG(BodyLong, i - 2) && US(i - 2) < AVG(ShadowShort, i - 2) &&
          ((RB(i - 1) < RB(i - 2) - AVG(Far, i - 2) && RB(i) < RB(i - 1) + AVG(Near, i - 1)) ||
           (RB(i) < RB(i - 1) - AVG(Far, i - 1)) ||
           (RB(i) < RB(i - 1) && RB(i - 1) < RB(i - 2) &&
            (US(i) > AVG(ShadowShort, i) || US(i - 1) > AVG(ShadowShort, i - 1))) ||
           (RB(i) < RB(i - 1) && US(i) > AVG(ShadowLong, i)))) ? -100 : 0;
}
P(stalledpattern) {
  return (COL(i - 2) == 1 && COL(i - 1) == 1 && COL(i) == 1 && C(i) > C(i - 1) &&
          C(i - 1) > C(i - 2) && RB(i - 2) > AVG(BodyLong, i - 2) &&
          RB(i - 1) > AVG(BodyLong, i - 1) && US(i - 1) < AVG(ShadowVeryShort, i - 1) &&
          O(i - 1) > O(i - 2) && O(i - 1) <= C(i - 2) + AVG(Near, i - 2) &&
          RB(i) < AVG(BodyShort, i) && O(i) >= C(i - 1) - RB(i) - AVG(Near, i - 1)) ? -100 : 0;
}
P(mathold) {
  const double pen = 0.5;
  return (RB(i - 4) > AVG(BodyLong, i - 4) && RB(i - 3) < AVG(BodyShort, i - 3) &&
          RB(i - 2) < AVG(BodyShort, i - 2) && RB(i - 1) < AVG(BodyShort, i - 1) &&
          COL(i - 4) == 1 && COL(i - 3) == -1 && COL(i) == 1 && RBGAPUP(i - 3, i - 4) &&
          BOT(i - 2) < C(i - 4) && BOT(i - 1) < C(i - 4) &&
          BOT(i - 2) > C(i - 4) - RB(i - 4) * pen && BOT(i - 1) > C(i - 4) - RB(i - 4) * pen &&
          TOP(i - 2) < O(i - 3) && TOP(i - 1) < TOP(i - 2) && O(i) > C(i - 1) &&
          C(i) > fmax(fmax(H(i - 3), H(i - 2)), H(i - 1))) ? 100 : 0;
}
/* HIKKAKE: pattern at i, or the confirmation (close beyond the high/low of
 * the bar before the pattern) of the latest pattern within the last 3
 * bars, once per pattern (ta_CDLHIKKAKE.c state machine, stateless). */
static int hik_at(const Bars* B, int i) {
  if (H(i - 1) < H(i - 2) && L(i - 1) > L(i - 2) &&
      ((H(i) < H(i - 1) && L(i) < L(i - 1)) || (H(i) > H(i - 1) && L(i) > L(i - 1))))
    return H(i) < H(i - 1) ? 100 : -100;
  return 0;
}
static int hikmod_at(const Bars* B, int i) {
  if (H(i - 2) < H(i - 3) && L(i - 2) > L(i - 3) && H(i - 1) < H(i - 2) && L(i - 1) > L(i - 2) &&
      ((H(i) < H(i - 1) && L(i) < L(i - 1) && C(i - 2) <= L(i - 2) + AVG(Near, i - 2)) ||
       (H(i) > H(i - 1) && L(i) > L(i - 1) && C(i - 2) >= H(i - 2) - AVG(Near, i - 2))))
    return H(i) < H(i - 1) ? 100 : -100;
  return 0;
}
static int hik_conf(const Bars* B, int i, int mod) {
  int pat = mod ? hikmod_at(B, i) : hik_at(B, i);
  if (pat) return pat;
  const int lo = mod ? 3 : 2;
  for (int j = i - 1; j >= i - 3 && j >= lo; --j) {
    const int pj = mod ? hikmod_at(B, j) : hik_at(B, j);
    if (!pj) continue;
    for (int m = j + 1; m <= i; ++m) {
      const int conf = (pj > 0 && C(m) > H(j - 1)) || (pj < 0 && C(m) < L(j - 1));
      if (conf) return m == i ? pj + (pj > 0 ? 100 : -100) : 0;
    }
    return 0;
  }
  return 0;
}
P(hikkake) { return hik_conf(B, i, 0); }
P(hikkakemod) { return hik_conf(B, i, 1); }
P(engulfing) {
  const int white = COL(i) == 1 && COL(i - 1) == -1 &&
                    ((C(i) >= O(i - 1) && O(i) < C(i - 1)) || (C(i) > O(i - 1) && O(i) <= C(i - 1)));
  const int black = COL(i) == -1 && COL(i - 1) == 1 &&
                    ((O(i) >= C(i - 1) && C(i) < O(i - 1)) || (O(i) > C(i - 1) && C(i) <= O(i - 1)));
  if (white || black) return (O(i) != C(i - 1) && C(i) != O(i - 1)) ? COL(i) * 100 : COL(i) * 80;
  return 0;
}
P(threewhitesoldiers) {
  return (COL(i - 2) == 1 && US(i - 2) < AVG(ShadowVeryShort, i - 2) && COL(i - 1) == 1 &&
          US(i - 1) < AVG(ShadowVeryShort, i - 1) && COL(i) == 1 &&
          US(i) < AVG(ShadowVeryShort, i) && C(i) > C(i - 1) && C(i - 1) > C(i - 2) &&
          O(i - 1) > O(i - 2) && O(i - 1) <= C(i - 2) + AVG(Near, i - 2) && O(i) > O(i - 1) &&
          O(i) <= C(i - 1) + AVG(Near, i - 1) && RB(i - 1) > RB(i - 2) - AVG(Far, i - 2) &&
          RB(i) > RB(i - 1) - AVG(Far, i - 1) && RB(i) > AVG(BodyShort, i)) ? 100 : 0;
}
P(darkcloudcover) {
  const double pen = 0.5;
  return (COL(i - 1) == 1 && RB(i - 1) > AVG(BodyLong, i - 1) && COL(i) == -1 &&
          O(i) > H(i - 1) && C(i) > O(i - 1) && C(i) < C(i - 1) - RB(i - 1) * pen) ? -100 : 0;
}
P(dojistar) {
  if (RB(i - 1) > AVG(BodyLong, i - 1) && RB(i) <= AVG(BodyDoji, i) &&
      ((COL(i - 1) == 1 && RBGAPUP(i, i - 1)) || (COL(i - 1) == -1 && RBGAPDOWN(i, i - 1))))
    return -COL(i - 1) * 100;
  return 0;
}
static int star(const Bars* B, int i, int doji, int up) {
  const double pen = 0.3;
  const int mid = doji ? (RB(i - 1) <= AVG(BodyDoji, i - 1)) : (RB(i - 1) <= AVG(BodyShort, i - 1));
  if (up)  /* evening: white, gap up, black closing into the first body */
    return (RB(i - 2) > AVG(BodyLong, i - 2) && COL(i - 2) == 1 && mid && RBGAPUP(i - 1, i - 2) &&
            RB(i) > AVG(BodyShort, i) && COL(i) == -1 && C(i) < C(i - 2) - RB(i - 2) * pen)
               ? -100 : 0;
  return (RB(i - 2) > AVG(BodyLong, i - 2) && COL(i - 2) == -1 && mid && RBGAPDOWN(i - 1, i - 2) &&
          RB(i) > AVG(BodyShort, i) && COL(i) == 1 && C(i) > C(i - 2) + RB(i - 2) * pen)
             ? 100 : 0;
}
P(eveningdojistar) { return star(B, i, 1, 1); }
P(eveningstar) { return star(B, i, 0, 1); }
P(morningdojistar) { return star(B, i, 1, 0); }
P(morningstar) { return star(B, i, 0, 0); }
static int harami_(const Bars* B, int i, int cross) {
  const int small = cross ? (RB(i) <= AVG(BodyDoji, i)) : (RB(i) <= AVG(BodyShort, i));
  if (RB(i - 1) > AVG(BodyLong, i - 1) && small) {
    if (TOP(i) < TOP(i - 1) && BOT(i) > BOT(i - 1)) return -COL(i - 1) * 100;
    if (TOP(i) <= TOP(i - 1) && BOT(i) >= BOT(i - 1)) return -COL(i - 1) * 80;
  }
  return 0;
}
P(harami) { return harami_(B, i, 0); }
P(haramicross) { return harami_(B, i, 1); }
static int kick_common(const Bars* B, int i) {
  return COL(i - 1) == -COL(i) && RB(i - 1) > AVG(BodyLong, i - 1) &&
         US(i - 1) < AVG(ShadowVeryShort, i - 1) && LS(i - 1) < AVG(ShadowVeryShort, i - 1) &&
         RB(i) > AVG(BodyLong, i) && US(i) < AVG(ShadowVeryShort, i) &&
         LS(i) < AVG(ShadowVeryShort, i) &&
         ((COL(i - 1) == -1 && GAPUP(i, i - 1)) || (COL(i - 1) == 1 && GAPDOWN(i, i - 1)));
}
P(kicking) { return kick_common(B, i) ? COL(i) * 100 : 0; }
P(kickingbylength) {
  return kick_common(B, i) ? COL(RB(i) > RB(i - 1) ? i : i - 1) * 100 : 0;
}
P(identical3crows) {
  return (COL(i - 2) == -1 && LS(i - 2) < AVG(ShadowVeryShort, i - 2) && COL(i - 1) == -1 &&
          LS(i - 1) < AVG(ShadowVeryShort, i - 1) && COL(i) == -1 &&
          LS(i) < AVG(ShadowVeryShort, i) && C(i - 2) > C(i - 1) && C(i - 1) > C(i) &&
          O(i - 1) <= C(i - 2) + AVG(Equal, i - 2) && O(i - 1) >= C(i - 2) - AVG(Equal, i - 2) &&
          O(i) <= C(i - 1) + AVG(Equal, i - 1) && O(i) >= C(i - 1) - AVG(Equal, i - 1)) ? -100 : 0;
}
P(ladderbottom) {
  return (COL(i - 4) == -1 && COL(i - 3) == -1 && COL(i - 2) == -1 && O(i - 4) > O(i - 3) &&
          O(i - 3) > O(i - 2) && C(i - 4) > C(i - 3) && C(i - 3) > C(i - 2) && COL(i - 1) == -1 &&
          US(i - 1) > AVG(ShadowVeryShort, i - 1) && COL(i) == 1 && O(i) > O(i - 1) &&
          C(i) > H(i - 1)) ? 100 : 0;
}
P(piercing) {
  return (COL(i - 1) == -1 && RB(i - 1) > AVG(BodyLong, i - 1) && COL(i) == 1 &&
          RB(i) > AVG(BodyLong, i) && O(i) < L(i - 1) && C(i) < O(i - 1) &&
          C(i) > C(i - 1) + RB(i - 1) * 0.5) ? 100 : 0;
}
P(sticksandwich) {
  return (COL(i - 2) == -1 && COL(i - 1) == 1 && COL(i) == -1 && L(i - 1) > C(i - 2) &&
          C(i) <= C(i - 2) + AVG(Equal, i - 2) && C(i) >= C(i - 2) - AVG(Equal, i - 2)) ? 100 : 0;
}
P(tasukigap) {
  if ((RBGAPUP(i - 1, i - 2) && COL(i - 1) == 1 && COL(i) == -1 && O(i) < C(i - 1) &&
       O(i) > O(i - 1) && C(i) < O(i - 1) && C(i) > TOP(i - 2) &&
       fabs(RB(i - 1) - RB(i)) < AVG(Near, i - 1)) ||
      (RBGAPDOWN(i - 1, i - 2) && COL(i - 1) == -1 && COL(i) == 1 && O(i) < O(i - 1) &&
       O(i) > C(i - 1) && C(i) > O(i - 1) && C(i) < BOT(i - 2) &&
       fabs(RB(i - 1) - RB(i)) < AVG(Near, i - 1)))
    return COL(i - 1) * 100;
  return 0;
}
P(threelinestrike) {
  if (COL(i - 3) == COL(i - 2) && COL(i - 2) == COL(i - 1) && COL(i) == -COL(i - 1) &&
      O(i - 2) >= BOT(i - 3) - AVG(Near, i - 3) && O(i - 2) <= TOP(i - 3) + AVG(Near, i - 3) &&
      O(i - 1) >= BOT(i - 2) - AVG(Near, i - 2) && O(i - 1) <= TOP(i - 2) + AVG(Near, i - 2) &&
      ((COL(i - 1) == 1 && C(i - 1) > C(i - 2) && C(i - 2) > C(i - 3) && O(i) > C(i - 1) &&
        C(i) < O(i - 3)) ||
       (COL(i - 1) == -1 && C(i - 1) < C(i - 2) && C(i - 2) < C(i - 3) && O(i) < C(i - 1) &&
        C(i) > O(i - 3))))
    return COL(i - 1) * 100;
  return 0;
}
P(threeoutside) {
  if ((COL(i - 1) == 1 && COL(i - 2) == -1 && C(i - 1) > O(i - 2) && O(i - 1) < C(i - 2) &&
       C(i) > C(i - 1)) ||
      (COL(i - 1) == -1 && COL(i - 2) == 1 && O(i - 1) > C(i - 2) && C(i - 1) < O(i - 2) &&
       C(i) < C(i - 1)))
    return COL(i - 1) * 100;
  return 0;
}
P(breakaway) {
  if (RB(i - 4) > AVG(BodyLong, i - 4) && COL(i - 4) == COL(i - 3) && COL(i - 3) == COL(i - 1) &&
      COL(i - 1) == -COL(i) &&
      ((COL(i - 4) == -1 && RBGAPDOWN(i - 3, i - 4) && H(i - 2) < H(i - 3) && L(i - 2) < L(i - 3) &&
        H(i - 1) < H(i - 2) && L(i - 1) < L(i - 2) && C(i) > O(i - 3) && C(i) < C(i - 4)) ||
       (COL(i - 4) == 1 && RBGAPUP(i - 3, i - 4) && H(i - 2) > H(i - 3) && L(i - 2) > L(i - 3) &&
        H(i - 1) > H(i - 2) && L(i - 1) > L(i - 2) && C(i) < O(i - 3) && C(i) > C(i - 4))))
    return COL(i) * 100;
  return 0;
}
P(twocrows) {
  return (COL(i - 2) == 1 && RB(i - 2) > AVG(BodyLong, i - 2) && COL(i - 1) == -1 &&
          RBGAPUP(i - 1, i - 2) && COL(i) == -1 && O(i) < O(i - 1) && O(i) > C(i - 1) &&
          C(i) > O(i - 2) && C(i) < C(i - 2)) ? -100 : 0;
}
P(unique3river) {
  return (RB(i - 2) > AVG(BodyLong, i - 2) && COL(i - 2) == -1 && COL(i - 1) == -1 &&
          C(i - 1) > C(i - 2) && O(i - 1) <= O(i - 2) && L(i - 1) < L(i - 2) &&
          RB(i) < AVG(BodyShort, i) && COL(i) == 1 && O(i) > L(i - 1)) ? 100 : 0;
}
P(xsidegap3methods) {
  if (COL(i - 2) == COL(i - 1) && COL(i - 1) == -COL(i) && O(i) < TOP(i - 1) &&
      O(i) > BOT(i - 1) && C(i) < TOP(i - 2) && C(i) > BOT(i - 2) &&
      ((COL(i - 2) == 1 && RBGAPUP(i - 1, i - 2)) || (COL(i - 2) == -1 && RBGAPDOWN(i - 1, i - 2))))
    return COL(i - 2) * 100;
  return 0;
}
P(risefall3methods) {
  const int c4 = COL(i - 4);
  return (RB(i - 4) > AVG(BodyLong, i - 4) && RB(i - 3) < AVG(BodyShort, i - 3) &&
          RB(i - 2) < AVG(BodyShort, i - 2) && RB(i - 1) < AVG(BodyShort, i - 1) &&
          RB(i) > AVG(BodyLong, i) && c4 == -COL(i - 3) && COL(i - 3) == COL(i - 2) &&
          COL(i - 2) == COL(i - 1) && COL(i - 1) == -COL(i) && BOT(i - 3) < H(i - 4) &&
          TOP(i - 3) > L(i - 4) && BOT(i - 2) < H(i - 4) && TOP(i - 2) > L(i - 4) &&
          BOT(i - 1) < H(i - 4) && TOP(i - 1) > L(i - 4) && C(i - 2) * c4 < C(i - 3) * c4 &&
          C(i - 1) * c4 < C(i - 2) * c4 && O(i) * c4 > C(i - 1) * c4 &&
          C(i) * c4 > C(i - 4) * c4) ? 100 * c4 : 0;
}

typedef int (*Rule)(const Bars*, int);
typedef struct {
  const char* name; /* the reference's ct_cdl_<name>_d suffix */
  int lookback;
  Rule fn;
} Pattern;

/* The 61 TA-Lib patterns in a fixed order (the GPU kernel's order). */
static const Pattern PATTERNS[] = {
    {"two_crows", 12, twocrows},
    {"three_black_crows", 13, threeblackcrows},
    {"three_inside", 12, threeinside},
    {"three_line_strike", 8, threelinestrike},
    {"three_outside", 3, threeoutside},
    {"three_stars_in_south", 12, threestarsinsouth},
    {"three_white_soldiers", 12, threewhitesoldiers},
    {"abandoned_baby", 12, abandonedbaby},
    {"advance_block", 12, advanceblock},
    {"belt_hold", 10, belthold},
    {"breakaway", 14, breakaway},
    {"closing_marubozu", 10, closingmarubozu},
    {"conceal_baby_swallow", 13, concealbabyswall},
    {"counterattack", 11, counterattack},
    {"dark_cloud_cover", 11, darkcloudcover},
    {"doji", 10, doji},
    {"doji_star", 11, dojistar},
    {"dragonfly_doji", 10, dragonflydoji},
    {"engulfing", 2, engulfing},
    {"evening_doji_star", 12, eveningdojistar},
    {"evening_star", 12, eveningstar},
    {"gap_side_side_white", 7, gapsidesidewhite},
    {"gravestone_doji", 10, gravestonedoji},
    {"hammer", 11, hammer},
    {"hanging_man", 11, hangingman},
    {"harami", 11, harami},
    {"harami_cross", 11, haramicross},
    {"high_wave", 10, highwave},
    {"hikkake", 5, hikkake},
    {"hikkake_mod", 10, hikkakemod},
    {"homing_pigeon", 11, homingpigeon},
    {"identical_three_crows", 12, identical3crows},
    {"in_neck", 11, inneck},
    {"inverted_hammer", 11, invertedhammer},
    {"kicking", 11, kicking},
    {"kicking_by_length", 11, kickingbylength},
    {"ladder_bottom", 14, ladderbottom},
    {"long_legged_doji", 10, longleggeddoji},
    {"long_line", 10, longline},
    {"marubozu", 10, marubozu},
    {"matching_low", 6, matchinglow},
    {"mat_hold", 14, mathold},
    {"morning_doji_star", 12, morningdojistar},
    {"morning_star", 12, morningstar},
    {"on_neck", 11, onneck},
    {"piercing", 11, piercing},
    {"rickshaw_man", 10, rickshawman},
    {"rise_fall3_methods", 14, risefall3methods},
    {"separating_lines", 11, separatinglines},
    {"shooting_star", 11, shootingstar},
    {"short_line", 10, shortline},
    {"spinning_top", 10, spinningtop},
    {"stalled_pattern", 12, stalledpattern},
    {"stick_sandwich", 7, sticksandwich},
    {"takuri", 10, takuri},
    {"tasuki_gap", 7, tasukigap},
    {"thrusting", 11, thrusting},
    {"tristar", 12, tristar},
    {"unique_3_river", 12, unique3river},
    {"upside_gap_2_crows", 12, upsidegap2crows},
    {"xside_gap_3_methods", 2, xsidegap3methods},
};
#define NPAT ((int)(sizeof PATTERNS / sizeof PATTERNS[0]))

API int orc_cdl_count(void) { return NPAT; }
API const char* orc_cdl_name(int k) { return (k >= 0 && k < NPAT) ? PATTERNS[k].name : 0; }
API int orc_cdl_lookback(int k) { return (k >= 0 && k < NPAT) ? PATTERNS[k].lookback : -1; }

API int orc_cdl_index(const float* o, const float* h, const float* l, const float* c,
                      int32_t* out, int n, int k) {
  if (!o || !h || !l || !c || !out || n <= 0 || k < 0 || k >= NPAT) return 1;
  const Bars B = {o, h, l, c};
  const Pattern* p = &PATTERNS[k];
  for (int i = 0; i < n; ++i) out[i] = i < p->lookback ? 0 : p->fn(&B, i);
  return 0;
}

API int orc_cdl(const char* name, const float* o, const float* h, const float* l,
                const float* c, int32_t* out, int n) {
  for (int k = 0; k < NPAT; ++k)
    if (name && strcmp(name, PATTERNS[k].name) == 0) return orc_cdl_index(o, h, l, c, out, n, k);
  return 1;
}
