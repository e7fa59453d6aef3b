// ddx core, part 3: cheap trick bounds and move-ordering weights.
//
// These are the DDS heuristics (QuickTricks, QuickTricksSecondHand, LaterTricksMIN/MAX and
// the lead / 2nd / 3rd / 4th hand sorting weights) by Bo Haglund and Soren Hein
// (https://github.com/dds-bridge/dds, Apache-2.0), as used by the reference solver
// (src/dd_bounds.h). The decision rules and weight tables must stay exactly the same as the
// reference's so that both solvers search identical trees; the code is restated over packed
// register data (suit lengths as nibbles, top-card holders and cards as bit fields) so that
// none of it needs per-lane scratch memory on the GPU.
//
// Conventions: "MAX" is the NS side (even seats), "MIN" EW. `wr` ("win ranks" in DDS) is the
// card mask a bound's answer depends on; it becomes part of the node's relevant cards.
#pragma once
#include "bits.hpp"

namespace ddx {

DDX_FN int partner(int h) { return (h + 2) & 3; }
DDX_FN int lho(int h) { return (h + 1) & 3; }
DDX_FN int rho(int h) { return (h + 3) & 3; }
DDX_FN bool is_ns(int h) { return (h & 1) == 0; }

// Suit lengths and the three highest cards of every suit with their holders
// (reference DDTop / dd_top_init, src/dd_bounds.h:12-32). A missing card reads as card 0
// held by seat -1, as in the reference.
struct Top {
  u64 lens;  // lengths() of the hands
  u64 hold;  // holder of the k-th highest card of suit s in nibble 3s+k; 0xF = none
  u32 c0, c1, c2;  // byte s of c0 / c1 / c2: the highest / second / third card left in suit s
};

DDX_FN int tlen(const Top& t, int h, int s) { return len_of(t.lens, h, s); }
DDX_FN int holder(const Top& t, int s, int k) {
  int v = (int)((t.hold >> (4 * (3 * s + k))) & 15);
  return v == 15 ? -1 : v;
}
DDX_FN int win_h(const Top& t, int s) { return holder(t, s, 0); }
DDX_FN int sec_h(const Top& t, int s) { return holder(t, s, 1); }
DDX_FN int thr_h(const Top& t, int s) { return holder(t, s, 2); }
DDX_FN int win_c(const Top& t, int s) { return (int)((t.c0 >> (8 * s)) & 0xFF); }
DDX_FN int sec_c(const Top& t, int s) { return (int)((t.c1 >> (8 * s)) & 0xFF); }
DDX_FN int thr_c(const Top& t, int s) { return (int)((t.c2 >> (8 * s)) & 0xFF); }

DDX_FN void top_init(Top& t, const u64* h, u64 lens) {
  t.lens = lens;
  t.hold = 0;
  t.c0 = t.c1 = t.c2 = 0;
  u64 all = h[0] | h[1] | h[2] | h[3];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    u64 a = all & suit_mask(s);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      int card = a ? msb(a) : 0;
      u64 who = a ? (u64)owner(h, card) : 15;
      t.hold |= who << (4 * (3 * s + k));
      u32 cb = (u32)card << (8 * s);
      if (k == 0) t.c0 |= cb;
      if (k == 1) t.c1 |= cb;
      if (k == 2) t.c2 |= cb;
      a &= ~bit(card);
    }
  }
}

// Suit order of the QuickTricks scan: trump first, then the others upward.
DDX_FN int qt_next_suit(int suit, int trump) {
  if (trump >= 0 && suit == trump) return trump == 0 ? 1 : 0;
  suit++;
  if (trump >= 0 && suit == trump) suit++;
  return suit;
}

// Outcome of one suit's quick-trick count.
enum { QT_SAME = 0, QT_CUT = 1, QT_NEXT = 2 };

// The leading hand holds the top card of `suit` (a side suit in a trump contract): its own
// winners, cashed before the opponents can ruff (DDS QtricksLeadHandTrump).
DDX_FN int qt_own_side_suit(const Top& t, int hand, int cutoff, int cL, int cR, int lhoT, int rhoT,
                            int cO, int cP, int suit, int qt, u64& wr, int& res) {
  res = QT_CUT;
  if ((cL != 0 || lhoT == 0) && (cR != 0 || rhoT == 0)) {
    wr |= bit(win_c(t, suit));
    if (++qt >= cutoff) return qt;
    if (cL <= 1 && cR <= 1 && cP <= 1 && lhoT == 0 && rhoT == 0) {
      qt += cO - 1;
      if (qt >= cutoff) return qt;
      res = QT_NEXT;
      return qt;
    }
  }
  int sh = sec_h(t, suit);
  if (sh == hand) {
    if (lhoT == 0 && rhoT == 0) {
      wr |= bit(sec_c(t, suit));
      if (++qt >= cutoff) return qt;
      if (cL <= 2 && cR <= 2 && cP <= 2) {
        qt += cO - 2;
        if (qt >= cutoff) return qt;
        res = QT_NEXT;
        return qt;
      }
    }
  } else if (sh == partner(hand) && cO > 1 && cP > 1) {
    if (lhoT == 0 && rhoT == 0) {
      wr |= bit(sec_c(t, suit));
      if (++qt >= cutoff) return qt;
      if (cL <= 2 && cR <= 2 && (cP <= 2 || cO <= 2)) {
        qt += (cO > cP ? cO : cP) - 2;
        if (qt >= cutoff) return qt;
        res = QT_NEXT;
        return qt;
      }
    }
  }
  res = QT_SAME;
  return qt;
}

// Cashing one trump drops a trump of each opponent (in the trump suit itself, unless it is
// the suit used to reach partner).
DDX_FN void qt_draw(int trump, int suit, bool comm, int comm_suit, int& lhoT, int& rhoT) {
  if (trump == suit && (!comm || suit != comm_suit)) {
    lhoT = lhoT > 0 ? lhoT - 1 : 0;
    rhoT = rhoT > 0 ? rhoT - 1 : 0;
  }
}

// The leading hand holds the top card of `suit`, in notrump or in the trump suit itself
// (DDS QtricksLeadHandNT).
DDX_FN int qt_own_suit(const Top& t, int hand, int cutoff, int cL, int cR, int& lhoT, int& rhoT,
                       bool comm, int comm_suit, int cO, int cP, int suit, int qt, int trump,
                       u64& wr, int& res) {
  res = QT_CUT;
  wr |= bit(win_c(t, suit));
  if (++qt >= cutoff) return qt;
  qt_draw(trump, suit, comm, comm_suit, lhoT, rhoT);
  if (cL <= 1 && cR <= 1 && cP <= 1) {
    qt += cO - 1;
    if (qt >= cutoff) return qt;
    res = QT_NEXT;
    return qt;
  }
  int sh = sec_h(t, suit);
  if (sh == hand) {
    wr |= bit(sec_c(t, suit));
    if (++qt >= cutoff) return qt;
    qt_draw(trump, suit, comm, comm_suit, lhoT, rhoT);
    if (cL <= 2 && cR <= 2 && cP <= 2) {
      qt += cO - 2;
      if (qt >= cutoff) return qt;
      res = QT_NEXT;
      return qt;
    }
  } else if (sh == partner(hand) && cO > 1 && cP > 1) {
    wr |= bit(sec_c(t, suit));
    if (++qt >= cutoff) return qt;
    qt_draw(trump, suit, comm, comm_suit, lhoT, rhoT);
    if (cL <= 2 && cR <= 2 && (cP <= 2 || cO <= 2)) {
      qt += (cO > cP ? cO : cP) - 2;
      if (qt >= cutoff) return qt;
      res = QT_NEXT;
      return qt;
    }
  }
  res = QT_SAME;
  return qt;
}

// Partner holds the top card of a side suit, reached through the entry card comm_card
// (DDS QtricksLeadHandTrump, partner branch).
DDX_FN int qt_partner_side_suit(const Top& t, int hand, int cutoff, int cL, int cR, int lhoT,
                                int rhoT, int cO, int cP, int suit, int qt, int comm_suit,
                                int comm_card, u64& wr, int& res) {
  res = QT_CUT;
  int P = partner(hand);
  if ((cL != 0 || lhoT == 0) && (cR != 0 || rhoT == 0)) {
    wr |= bit(win_c(t, suit)) | bit(comm_card);
    if (++qt >= cutoff) return qt;
    if (cL <= 1 && cR <= 1 && cO <= 1 && lhoT == 0 && rhoT == 0) {
      qt += cP - 1;
      if (qt >= cutoff) return qt;
      res = QT_NEXT;
      return qt;
    }
  }
  int sh = sec_h(t, suit);
  if (sh == P) {
    if (lhoT == 0 && rhoT == 0) {
      wr |= bit(sec_c(t, suit)) | bit(comm_card);
      if (++qt >= cutoff) return qt;
      if (cL <= 2 && cR <= 2 && cO <= 2) {
        qt += cP - 2;
        if (qt >= cutoff) return qt;
        res = QT_NEXT;
        return qt;
      }
    }
  } else if (sh == hand && cP > 1 && cO > 1) {
    if (lhoT == 0 && rhoT == 0) {
      wr |= bit(sec_c(t, suit)) | bit(comm_card);
      if (++qt >= cutoff) return qt;
      if (cL <= 2 && cR <= 2 && (cO <= 2 || cP <= 2)) {
        qt += (cP > cO ? cP : cO) - 2;
        if (qt >= cutoff) return qt;
        res = QT_NEXT;
        return qt;
      }
    }
  } else if (suit == comm_suit && sh == lho(hand) && (cL >= 2 || lhoT == 0) &&
             (cR >= 2 || rhoT == 0)) {
    if (thr_h(t, suit) == P) {
      wr |= bit(thr_c(t, suit)) | bit(comm_card);
      if (++qt >= cutoff) return qt;
      if (cO <= 2 && cL <= 2 && cR <= 2 && lhoT == 0 && rhoT == 0) {
        qt += cP - 2;
        if (qt >= cutoff) return qt;
      }
    }
  }
  res = QT_SAME;
  return qt;
}

// Partner holds the top card, notrump or the trump suit (DDS QtricksLeadHandNT, partner).
DDX_FN int qt_partner_suit(const Top& t, int hand, int cutoff, int cL, int cR, int cO, int cP,
                           int suit, int qt, int comm_suit, int comm_card, u64& wr, int& res) {
  res = QT_CUT;
  int P = partner(hand);
  wr |= bit(win_c(t, suit)) | bit(comm_card);
  if (++qt >= cutoff) return qt;
  if (cL <= 1 && cR <= 1 && cO <= 1) {
    qt += cP - 1;
    if (qt >= cutoff) return qt;
    res = QT_NEXT;
    return qt;
  }
  int sh = sec_h(t, suit);
  if (sh == P) {
    wr |= bit(sec_c(t, suit));
    if (++qt >= cutoff) return qt;
    if (cL <= 2 && cR <= 2 && cO <= 2) {
      qt += cP - 2;
      if (qt >= cutoff) return qt;
      res = QT_NEXT;
      return qt;
    }
  } else if (sh == hand && cP > 1 && cO > 1) {
    wr |= bit(sec_c(t, suit));
    if (++qt >= cutoff) return qt;
    if (cL <= 2 && cR <= 2 && (cO <= 2 || cP <= 2)) {
      qt += (cP > cO ? cP : cO) - 2;
      if (qt >= cutoff) return qt;
      res = QT_NEXT;
      return qt;
    }
  } else if (suit == comm_suit && sh == lho(hand)) {
    if (thr_h(t, suit) == P) {
      wr |= bit(thr_c(t, suit));
      if (++qt >= cutoff) return qt;
      if (cO <= 2 && cL <= 2 && cR <= 2) {
        qt += cP - 2;
        if (qt >= cutoff) return qt;
      }
    }
  }
  res = QT_SAME;
  return qt;
}

// Is there a side suit s (not trump) in which `a` has cards and `b` none? (a ruffing entry)
DDX_FN bool has_void_entry(const Top& t, int trump, int a, int b) {
  bool any = false;
#pragma unroll
  for (int s = 0; s < 4; s++) any = any || (s != trump && tlen(t, a, s) > 0 && tlen(t, b, s) == 0);
  return any;
}

// Quick-trick bound of DDS for the leader's side (`hand` leads). The leader's side is short
// of `cutoff` tricks, its opponents of `cutoff_other`. Returns the quick tricks;
// decided = true when that settles the node: then a positive count means the leading side
// makes it, 0 that it fails (reference dd_quick_tricks, src/dd_bounds.h:245-438).
DDX_FN int quick_tricks(const u64* h, const Top& t, int trump, int hand, int cutoff,
                        int cutoff_other, u64& wr, bool& decided) {
  const int P = partner(hand), L = lho(hand), R = rho(hand);
  int comm_card = 0, comm_suit = -1, ruffed = 0, qt = 0;
  bool comm = false;
  decided = true;

  // An entry to partner: partner's winner in a suit the leader can lead (in trump
  // contracts one the opponents cannot ruff), or a second-round winner shared by both.
  for (int s = 0; s < 4 && !comm; s++) {
    if (trump >= 0 && s == trump) continue;
    bool safe = trump < 0 || ((tlen(t, L, s) != 0 || tlen(t, L, trump) == 0) &&
                              (tlen(t, R, s) != 0 || tlen(t, R, trump) == 0));
    if (win_h(t, s) == P) {
      if (tlen(t, hand, s) != 0 && safe) {
        comm = true;
        comm_suit = s;
        comm_card = win_c(t, s);
      }
    } else if (sec_h(t, s) == P && win_h(t, s) == hand && tlen(t, hand, s) >= 2 &&
               tlen(t, P, s) >= 2 && safe) {
      comm = true;
      comm_suit = s;
      comm_card = sec_c(t, s);
    }
  }
  if (trump >= 0 && !comm && tlen(t, hand, trump) != 0 && win_h(t, trump) == P) {
    comm = true;
    comm_suit = trump;
    comm_card = win_c(t, trump);
  }

  int lhoT = 0, rhoT = 0, suit = 0;
  if (trump >= 0) {
    suit = trump;
    lhoT = tlen(t, L, trump);
    rhoT = tlen(t, R, trump);
  }
  while (suit <= 3) {
    const int cO = tlen(t, hand, suit), cL = tlen(t, L, suit), cR = tlen(t, R, suit);
    const int cP = tlen(t, P, suit);
    const int opps = cL | cR;
    const bool side = trump >= 0 && trump != suit;  // a side suit of a trump contract
    if (!opps && cP == 0) {
      // Only the leader has the suit: all its cards are tricks (side suits only when the
      // opponents have no trumps left).
      if (cO != 0 && (!side || (lhoT == 0 && rhoT == 0))) {
        qt += cO;
        if (qt >= cutoff) return qt;
      }
      suit = qt_next_suit(suit, trump);
      continue;
    }
    // The opponents are void: the partnership's long trumps (plus one ruff), or the
    // shorter holding of a side suit the opponents cannot ruff.
    if (!opps && trump >= 0 && suit == trump) {
      int sum = cO > cP ? cO : cP;
      if (sum > 0 && cO >= cP && has_void_entry(t, trump, hand, P)) sum++;
      if (sum >= cutoff) return sum;
    } else if (!opps && (trump < 0 || (side && lhoT == 0 && rhoT == 0))) {
      int sum = cO < cP ? cO : cP;
      if (sum >= cutoff) return sum;
    }
    if (comm) {
      if (!opps && cO == 0) {
        // Only partner has the suit: cash it after the entry.
        if (!side || (lhoT == 0 && rhoT == 0)) {
          qt += cP;
          wr |= bit(comm_card);
          if (qt >= cutoff) return qt;
        }
        suit = qt_next_suit(suit, trump);
        continue;
      }
      if (!opps && trump >= 0 && suit == trump) {
        int sum = cO > cP ? cO : cP;
        if (sum > 0 && cO <= cP && has_void_entry(t, trump, P, hand)) sum++;
        if (sum >= cutoff) {
          wr |= bit(comm_card);
          return sum;
        }
      } else if (!opps && (trump < 0 || (side && lhoT == 0 && rhoT == 0))) {
        int sum = cO < cP ? cO : cP;
        if (sum >= cutoff) return sum;
      }
    }
    const int wh = win_h(t, suit);
    if (wh < 0) {
      suit = qt_next_suit(suit, trump);
      continue;
    }
    int res = QT_SAME;
    if (wh == hand) {
      qt = side ? qt_own_side_suit(t, hand, cutoff, cL, cR, lhoT, rhoT, cO, cP, suit, qt, wr, res)
                : qt_own_suit(t, hand, cutoff, cL, cR, lhoT, rhoT, comm, comm_suit, cO, cP, suit, qt,
                              trump, wr, res);
    } else if (wh == P && comm) {
      qt = side ? qt_partner_side_suit(t, hand, cutoff, cL, cR, lhoT, rhoT, cO, cP, suit, qt,
                                       comm_suit, comm_card, wr, res)
                : qt_partner_suit(t, hand, cutoff, cL, cR, cO, cP, suit, qt, comm_suit, comm_card,
                                  wr, res);
    }
    if (res == QT_CUT) return qt;
    if (res == QT_NEXT) {
      suit = qt_next_suit(suit, trump);
      continue;
    }

    // One ruff by partner, when the leader can lead the suit and partner is void in it.
    if (side && cO > 0 && ruffed == 0 &&
        (qt == 0 || (wh != hand && wh != P && win_h(t, trump) != hand && win_h(t, trump) != P)) &&
        cP == 0 && tlen(t, P, trump) > 0) {
      u64 T = suit_mask(trump), pt = seat_hand(h, P) & T;
      if ((cR > 0 || tlen(t, R, trump) == 0) && (cL > 0 || tlen(t, L, trump) == 0)) {
        ruffed = 1;
        if (1 >= cutoff) return 1;
        suit = qt_next_suit(suit, trump);
        continue;
      }
      // Some opponent can overruff: partner's top trump must beat theirs.
      const u64 hL = seat_hand(h, L), hR = seat_hand(h, R);
      u64 over = cR == 0 && cL == 0 ? (hL | hR) & T : cL == 0 ? hL & T : hR & T;
      if (over < pt) {
        ruffed = 1;
        wr |= bit(msb(pt));
        if (1 >= cutoff) return 1;
      }
      suit = qt_next_suit(suit, trump);
      continue;
    }
    if (qt >= cutoff) return qt;
    suit = qt_next_suit(suit, trump);
  }

  // Nothing cashes and trumps give no winner: whatever suit is led, a defender's top card
  // takes the trick.
  if (qt == 0 && (trump < 0 || win_h(t, trump) < 0)) {
#pragma unroll
    for (int s = 0; s < 4; s++)
      if (win_h(t, s) >= 0 && tlen(t, hand, s) > 0) wr = (wr & ~suit_mask(s)) | bit(win_c(t, s));
    if (1 >= cutoff_other) return 0;
  }
  decided = false;
  return qt;
}

// Second-hand quick tricks of DDS: after the opening card `lead`, does the partnership of the
// second player `hand` take this trick and, in notrump, reach `cutoff` tricks from there? `t` is built from the
// trick-start hands, `h` are the hands now (reference dd_quick_tricks_2nd,
// src/dd_bounds.h:442-479).
DDX_FN bool quick_tricks_2nd(const u64* h, const Top& t, int trump, int hand, int lead, int cutoff,
                             u64& wr) {
  const int ss = lead >> 4, P = partner(hand), L = lho(hand);
  const u64 S = suit_mask(ss);
  const u64 hO = seat_hand(h, hand), hP = seat_hand(h, P), hL = seat_hand(h, L);
  const u64 ranks = (hO | hP) & S;
  wr = 0;
  const bool side = trump >= 0 && ss != trump;
  const u64 T = trump >= 0 ? suit_mask(trump) : 0;
  const bool l_ruffs = !(hL & S) && (hL & T);
  if (side && ((!(hO & S) && (hO & T)) || (!(hP & S) && (hP & T)))) {
    if (l_ruffs) return false;  // our side ruffs, but the 4th hand can overruff
  } else if (ranks > (bit(lead) | (hL & S))) {
    if (side && l_ruffs) return false;
    wr = bit(msb(ranks));
  } else {
    return false;
  }
  int qt = 1;
  if (qt >= cutoff) return true;
  if (trump >= 0) return false;
  const bool own_hi = (hO & S) > (hP & S);
  const int hh = own_hi ? hand : P;
  const u64 hH = own_hi ? hO : hP, others = (h[0] | h[1] | h[2] | h[3]) & ~hH;
  if (win_h(t, ss) == hh && sec_h(t, ss) == hh) {
    wr |= bit(sec_c(t, ss));
    if (++qt >= cutoff) return true;
  }
#pragma unroll
  for (int s = 0; s < 4; s++) {
    if (s == ss || !(hH & suit_mask(s))) continue;
    if (!(others & suit_mask(s))) {
      qt += popc(hH & suit_mask(s));
      if (qt >= cutoff) return true;
    } else if (win_h(t, s) == hh) {
      wr |= bit(win_c(t, s));
      if (++qt >= cutoff) return true;
    }
  }
  return false;
}

// Tricks a side takes off the top in notrump: for each suit it holds the top card of, the
// longer of the two holdings (DDS LaterTricks, sum over top winners).
DDX_FN int top_suit_length_sum(const Top& t, bool ns_side) {
  int sum = 0;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    int hh = win_h(t, s);
    if (hh >= 0 && is_ns(hh) == ns_side) {
      int a = tlen(t, hh, s), b = tlen(t, partner(hh), s);
      sum += a > b ? a : b;
    }
  }
  return sum;
}
// Top cards of `ns_side` that the other hands still have to follow to.
DDX_FN u64 live_top_cards(const Top& t, bool ns_side) {
  u64 wr = 0;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    int wh = win_h(t, s);
    if (wh >= 0 && is_ns(wh) == ns_side &&
        (tlen(t, partner(wh), s) || tlen(t, lho(wh), s) || tlen(t, rho(wh), s)))
      wr |= bit(win_c(t, s));
  }
  return wr;
}
DDX_FN int longer_opp_trumps(const Top& t, int hand, int trump) {
  int a = tlen(t, lho(hand), trump), b = tlen(t, rho(hand), trump);
  return a > b ? a : b;
}

// DDS LaterTricksMIN, NS on lead (hand = the leader): false when NS cannot reach `target`
// (reference dd_later_min, src/dd_bounds.h:482-531).
DDX_FN bool later_min(const Top& t, int trump, int hand, int ns_won, int left, int target, u64& wr) {
  const int dq = left - 1;
  if (trump < 0 || win_h(t, trump) < 0) {
    int sum = top_suit_length_sum(t, true);
    if (ns_won + sum < target && sum > 0) {
      if (ns_won + dq >= target) return true;
      wr = live_top_cards(t, false);
      return false;
    }
  } else if (!is_ns(win_h(t, trump))) {
    if (tlen(t, hand, trump) == 0 && tlen(t, partner(hand), trump) == 0) {
      if (ns_won + dq + 1 - longer_opp_trumps(t, hand, trump) < target) {
        wr = 0;
        return false;
      }
    } else if (ns_won + dq < target) {
      wr = bit(win_c(t, trump));
      return false;
    } else if (ns_won + dq == target) {
      int hh = sec_h(t, trump);
      if (hh < 0) return true;
      if (!is_ns(hh) && (tlen(t, hh, trump) > 1 || tlen(t, partner(hh), trump) > 1)) {
        wr = bit(sec_c(t, trump));
        return false;
      }
    }
  } else {
    int hh = sec_h(t, trump);
    if (hh < 0) return true;
    if (is_ns(hh) || tlen(t, hh, trump) <= 1) return true;
    if (win_h(t, trump) == rho(hh)) {
      if (ns_won + dq < target) {
        wr = bit(sec_c(t, trump));
        return false;
      }
    } else {
      int h3 = thr_h(t, trump);
      if (h3 < 0) return true;
      if (!is_ns(h3) && ns_won + dq < target) {
        wr = bit(thr_c(t, trump));
        return false;
      }
    }
  }
  return true;
}

// DDS LaterTricksMAX, EW on lead: true when NS reach `target` anyway
// (reference dd_later_max, src/dd_bounds.h:534-584).
DDX_FN bool later_max(const Top& t, int trump, int hand, int ns_won, int left, int target, u64& wr) {
  const int dq = left - 1;
  if (trump < 0 || win_h(t, trump) < 0) {
    int sum = top_suit_length_sum(t, false);
    if (ns_won + dq + 1 - sum >= target && sum > 0) {
      if (ns_won + 1 < target) return false;
      wr = live_top_cards(t, true);
      return true;
    }
  } else if (is_ns(win_h(t, trump))) {
    if (tlen(t, hand, trump) == 0 && tlen(t, partner(hand), trump) == 0) {
      if (ns_won + longer_opp_trumps(t, hand, trump) >= target) {
        wr = 0;
        return true;
      }
    } else if (ns_won + 1 >= target) {
      wr = bit(win_c(t, trump));
      return true;
    } else {
      int hh = sec_h(t, trump);
      if (hh < 0) return false;
      if (is_ns(hh) && (tlen(t, hh, trump) > 1 || tlen(t, partner(hh), trump) > 1) &&
          ns_won + 2 >= target) {
        wr = bit(sec_c(t, trump));
        return true;
      }
    }
  } else {
    int hh = sec_h(t, trump);
    if (hh < 0) return false;
    if (!is_ns(hh) || tlen(t, hh, trump) <= 1) return false;
    if (win_h(t, trump) == rho(hh)) {
      if (ns_won + 1 >= target) {
        wr = bit(sec_c(t, trump));
        return true;
      }
    } else {
      int h3 = thr_h(t, trump);
      if (h3 < 0) return false;
      if (is_ns(h3) && ns_won + 1 >= target) {
        wr = bit(thr_c(t, trump));
        return true;
      }
    }
  }
  return false;
}

// ---------------------------------------------------------------------------------------
// Move ordering.

// DDS weight of leading `card` (weight_alloc_trump0 / weight_alloc_nt0); `tt` = the card or
// an equivalent one is the TT's stored best move (reference dd_lead_weight,
// src/dd_bounds.h:588-719).
DDX_FN int lead_weight(const u64* h, const Top& t, int trump, int left, int lead, int card, bool tt) {
  const int s = card >> 4, P = partner(lead), L = lho(lead), R = rho(lead);
  const u64 S = suit_mask(s), all = h[0] | h[1] | h[2] | h[3];
  const u64 hO = seat_hand(h, lead), hP = seat_hand(h, P), hL = seat_hand(h, L), hR = seat_hand(h, R);
  const int r_rank = popc(all & S & ~((1ull << card) - 1));  // 1 = the suit's top card
  const u64 below = all & S & ((1ull << card) - 1);
  const bool seq = below && ((hO >> msb(below)) & 1);
  const u64 risP = hP & S, risL = hL & S, risR = hR & S;
  const int lenL = tlen(t, L, s), lenR = tlen(t, R, s), lenP = tlen(t, P, s);
  const int cntLH = (lenL == 0 ? left : lenL) << 2;
  const int cntRH = (lenR == 0 ? left : lenR) << 2;
  const int wh = win_h(t, s), sh = sec_h(t, s), th3 = thr_h(t, s);
  const bool single_top_opp = (lenL == 1 && wh == L) || (lenR == 1 && wh == R);

  if (trump >= 0 && (all & suit_mask(trump))) {
    const u64 T = suit_mask(trump);
    const int tL = tlen(t, L, trump), tR = tlen(t, R, trump), tP = tlen(t, P, trump);
    int bonus = 0;
    if (s != trump && ((!risL && (hL & T)) || (!risR && (hR & T)))) bonus = -12;
    if (s != trump && lenP == 0 && tP > 0 && lenR > 0) bonus += 17;
    if (wh == R || sh == R) {
      if (lenR != 1) bonus += -12;
    } else if (wh == L && sh == P) {
      if (lenP != 1) bonus += 27;
    }
    if (s != trump && tlen(t, lead, s) == 1 && tlen(t, lead, trump) > 0 && lenP > 1 && wh == P)
      bonus += 19;
    int delta = bonus - (((cntLH + cntRH) << 5) / 13);
    bool win = false;
    const bool opp_cannot_ruff = (lenL != 0 || tL == 0) && (lenR != 0 || tR == 0);
    if (win_c(t, s) == card) {
      if (s != trump) {
        if (lenP != 0 || tP == 0) {
          win = opp_cannot_ruff;
        } else {
          win = (lenL != 0 || (hP & T) > (hL & T)) && (lenR != 0 || (hP & T) > (hR & T));
        }
      } else {
        win = true;
      }
    } else if (risP > (risL | risR)) {
      win = s != trump ? opp_cannot_ruff : true;
    } else if (s != trump && lenP == 0 && tP != 0) {
      const bool l_ruff = lenL == 0 && tL != 0, r_ruff = lenR == 0 && tR != 0;
      if (l_ruff && r_ruff) win = (hP & T) > ((hL | hR) & T);
      else if (l_ruff) win = (hP & T) > (hL & T);
      else if (r_ruff) win = (hP & T) > (hR & T);
      else win = true;
    }
    int w;
    if (win) {
      if (single_top_opp) w = delta + 35 + r_rank;
      else if (wh == lead) {
        if (sh == P) w = delta + 48 + r_rank;
        else if (win_c(t, s) == card) w = delta + 31;
        else w = delta - 3 + r_rank;
      } else if (wh == P) w = delta + (sh == lead ? 42 : 28) + r_rank;
      else if (seq && card == sec_c(t, s)) w = delta + 40;
      else if (seq) w = delta + 22 + r_rank;
      else w = delta + 11 + r_rank;
      if (tt) w += 18;
    } else {
      if (sh == P && P == th3) delta += 20;
      else if ((sh == lead && P == th3 && lenP > 1) || (sh == P && lead == th3 && lenP > 1))
        delta += 13;
      if (single_top_opp) w = delta + r_rank + 2;
      else if (wh == lead) {
        if (sh == P) w = delta + 33 + r_rank;
        else if (win_c(t, s) == card) w = delta + 38;
        else w = delta - 14 + r_rank;
      } else if (wh == P) w = delta + 34 + r_rank;
      else if (seq && card == sec_c(t, s)) w = delta + 35;
      else w = delta + 17 - ((card & 15) + 2);
    }
    return w;
  }

  // Notrump, or a trump contract with no trumps left.
  int delta = -(((cntLH + cntRH) << 5) / 19);
  if (lenP == 0) delta += -9;
  int w;
  if (win_c(t, s) == card || risP > (risL | risR)) {
    if (sh == R) {
      if (lenR != 1) delta += -1;
    } else if (sh == L) {
      delta += lenL != 1 ? 22 : 16;
    }
    if ((sh != L || lenL == 1) && (sh != R || lenR == 1)) w = delta + 45 + r_rank;
    else w = delta + 18 + r_rank;
    if (tt) w += 32;
  } else {
    if (wh == R || sh == R) {
      if (lenR != 1) delta += -10;
    } else if (wh == L && sh == P) {
      if (lenP != 1) delta += 31;
    }
    if (sh == P && P == th3) delta += 35;
    else if ((sh == lead && P == th3 && lenP > 1) || (sh == P && lead == th3 && lenP > 1))
      delta += 25;
    if (single_top_opp) w = delta + 28 + r_rank;
    else if (wh == lead) w = delta - 17 + r_rank;
    else if (!seq) w = delta + 12 + r_rank;
    else if (card == sec_c(t, s)) w = delta + 48;
    else w = delta + 29 - r_rank;
    if (tt) w += 19;
  }
  return w;
}

// Ranks as DDS numbers them (deuce = 2 .. ace = 14); an empty mask gives 0.
DDX_FN int rank_of(int card) { return (card & 15) + 2; }
DDX_FN int hi_rank(u64 m) { return m ? (msb(m) & 15) + 2 : 0; }
DDX_FN int lo_rank(u64 m) { return m ? (lsb(m) & 15) + 2 : 0; }

// Per-node context of the follow weights: the trick so far and who held the two top cards
// of each suit when the trick started (reference DDFollow / dd_follow_init,
// src/dd_bounds.h:734-763).
struct Follow {
  u64 allt;  // every card still out when the trick began
  u64 all;   // all cards now
  u32 top2;  // holder of the top / second card of suit s at trick start: nibbles 2s, 2s+1
  int lead, ls, lead0, high1, high2, move1, move2;
};

DDX_FN int follow_top2(const Follow& f, int s, int k) {
  int v = (int)((f.top2 >> (4 * (2 * s + k))) & 15);
  return v == 15 ? -1 : v;
}

// `h` = hands now, `trick` = packed cards played (byte i = i-th card), n of them.
DDX_FN void follow_init(Follow& f, const u64* h, int trump, int leader, u32 trick, int n) {
  u64 th[4] = {h[0], h[1], h[2], h[3]};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (i < n) {
      int card = (int)((trick >> (8 * i)) & 63);
      int seat = (leader + i) & 3;
#pragma unroll
      for (int p = 0; p < 4; p++)
        if (p == seat) th[p] |= bit(card);
    }
  }
  f.lead = leader;
  int c0 = (int)(trick & 63), c1 = (int)((trick >> 8) & 63), c2 = (int)((trick >> 16) & 63);
  f.ls = c0 >> 4;
  f.all = h[0] | h[1] | h[2] | h[3];
  f.allt = th[0] | th[1] | th[2] | th[3];
  u32 top2 = 0;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    u64 m = f.allt & suit_mask(s);
    u32 a = m ? (u32)owner(th, msb(m)) : 15;
    u64 m2 = m ? m ^ bit(msb(m)) : 0;
    u32 b = m2 ? (u32)owner(th, msb(m2)) : 15;
    top2 |= (a << (8 * s)) | (b << (8 * s + 4));
  }
  f.top2 = top2;
  f.lead0 = rank_of(c0);
  int w1 = n > 1 && beats(c1, c0, trump) ? 1 : 0;
  int cw1 = w1 ? c1 : c0;
  int w2 = n > 2 && beats(c2, cw1, trump) ? 2 : w1;
  f.high1 = w1;
  f.move1 = cw1;
  f.high2 = w2;
  f.move2 = w2 == 2 ? c2 : cw1;
}

// Move weights of DDS for the players after the lead (its weight_alloc_*1/2/3 rules; the
// rank_forces_ace and get_top_number refinements are not applied); n = cards already played in the trick
// (reference dd_follow_weight, src/dd_bounds.h:767-949).
DDX_FN int follow_weight(const Follow& f, const u64* h, int trump, int n, int seat, int card) {
  const int lead = f.lead, ls = f.ls;
  const u64 LS = suit_mask(ls);
  const int s = card >> 4, rank = rank_of(card);
  const int P = partner(lead), R = rho(lead);
  const u64 hS = seat_hand(h, seat), hP = seat_hand(h, P), hR = seat_hand(h, R), hO = seat_hand(h, lead);
  const int r_rank = popc(f.all & suit_mask(s) & ~((1ull << card) - 1));
  const u64 below = f.allt & suit_mask(s) & ((1ull << card) - 1);
  const bool seq = below && ((hS >> msb(below)) & 1);
  const int lead0 = f.lead0;
  const bool is_void = s != ls;
  const int cnt = popc(hS & suit_mask(s));
  const bool sec_mine = follow_top2(f, s, 1) == seat;
  const bool win_mine = follow_top2(f, s, 0) == seat;
  const u64 T = trump >= 0 ? suit_mask(trump) : 0;

  if (n == 1) {  // second hand; P = third hand, R = fourth hand
    const int max3rd = hi_rank(hP & LS), maxpd = hi_rank(hR & LS);
    const int min3rd = lo_rank(hP & LS), minpd = lo_rank(hR & LS);
    if (!is_void) {
      if (trump >= 0) {
        bool win = false;
        if (ls == trump) {
          win = (maxpd > lead0 && maxpd > max3rd) || (rank > lead0 && rank > max3rd);
        } else if (rank > lead0 && rank > max3rd) {
          win = (max3rd != 0 || !(hP & T)) ||
                (maxpd == 0 && (hR & T) && (hR & T) > (hP & T));
        } else if (maxpd > lead0 && maxpd > max3rd) {
          win = max3rd != 0 || !(hP & T);
        } else if (lead0 > maxpd && lead0 > max3rd && lead0 > rank) {
          win = maxpd == 0 && (hR & T) && ((max3rd != 0 || !(hP & T)) || (hR & T) > (hP & T));
        } else if (maxpd == 0 && (hR & T)) {
          win = true;
        }
        if (win) {
          if (min3rd > rank) return 40 + r_rank;
          if (maxpd > lead0 && (hO & LS) > (hR & LS)) return 41 + r_rank;
          if (rank > lead0) {
            if (rank < maxpd) return 78 - rank;
            if (rank > max3rd) return 73 - rank;
            return seq ? 62 - rank : 49 - rank;
          }
          return maxpd > 0 ? 47 - rank : 40 - rank;
        }
        if (rank < min3rd || rank < minpd) return -9 + r_rank;
        if (rank < lead0) return -16 + r_rank;
        return seq ? 22 - rank : 10 - rank;
      }
      if (maxpd > lead0 && maxpd > max3rd) return -rank;
      if (rank > lead0 && rank > max3rd) return 81 - rank;
      if (min3rd > rank || minpd > rank) return -3 + r_rank;
      if (rank < lead0) return -11 + r_rank;
      return seq ? 10 + r_rank : 13 - rank;
    }
    const u64 lead_bit = bit((ls << 4) | (lead0 - 2));
    const int add44 = (cnt << 6) / 44, add36 = (cnt << 6) / 36;
    const int short2 = cnt == 2 && sec_mine ? -4 : 0;
    if (trump >= 0) {
      if (ls == trump) {
        if ((hR & LS) > ((hP & LS) | lead_bit)) return -rank + add44;
        return -rank + add36 + short2;
      }
      if (s != trump) {
        bool hi;
        if (hP & LS) hi = (hR & LS) > ((hP & LS) | lead_bit) || (!(hR & LS) && (hR & T));
        else hi = (!(hR & LS) && (hR & T) > (hP & T)) || (!(hP & T) && (hR & LS) > lead_bit);
        return hi ? -rank + 60 + add44 : -rank - 2 + add36 + short2;
      }
      if ((hP & LS) || (!(hR & LS) && (hR & T) && (hR & T) > (hP & T)))
        return 24 - rank + add44;
      if (bit(card) > (hP & T)) return 24 - rank + add44;
      return 15 - rank + add36 + short2;
    }
    int add;
    if ((hR & LS) > ((hP & LS) | lead_bit)) {
      add = (cnt << 6) / 23;
      if (cnt == 2 && sec_mine) add += -2;
      else if (cnt == 1 && win_mine) add += -3;
    } else {
      add = (cnt << 6) / 33;
      if (cnt == 2 && sec_mine) add += -6;
      else if (cnt == 1 && win_mine) add += -8;
    }
    return -rank + add;
  }

  if (n == 2) {  // third hand; R = fourth hand
    const u64 c4 = hR & LS;
    const int max4th = hi_rank(c4), min4th = lo_rank(c4);
    const int m1_rank = rank_of(f.move1), m1_suit = f.move1 >> 4;
    if (!is_void) {
      const int max3rd = hi_rank(hS & LS);
      if (trump >= 0) {
        if (ls == trump) {
          if (f.high1 == 0 && lead0 > max4th) return -rank;
          if (max3rd < min4th || max3rd < m1_rank) return -rank;
          if (max3rd > max4th) return rank > max4th && rank > m1_rank ? 58 - rank : -rank;
          return -rank;
        }
        if (m1_suit == trump) return -rank;
        if (f.high1 == 0) {
          if (max4th == 0 || lead0 > max4th) return -rank;
          if (max3rd < min4th || max3rd < m1_rank) return -rank;
          if (max3rd > max4th) return rank > max4th ? 58 - rank : -rank;
          return rank > m1_rank && rank > max4th ? 60 - rank : -rank;
        }
        if (max4th == 0) return rank > m1_rank ? 20 - rank : -rank;
        if (max3rd < min4th || max3rd < m1_rank) return -rank;
        if (max3rd > max4th) return rank > m1_rank && rank > max4th ? 58 - rank : -rank;
        return rank > m1_rank && rank > max4th ? 60 - rank : -rank;
      }
      if (f.high1 == 0 && lead0 > max4th) return -rank;
      if (max3rd < min4th || max3rd < m1_rank) return -rank;
      return rank > m1_rank && rank > max4th ? 60 - rank : -rank;
    }
    if (trump >= 0) {
      if (ls == trump || s != trump) return -rank + (cnt << 6) / 40;
      if (f.high1 == 0 && lead0 > max4th && (max4th != 0 || !(hR & T))) return -rank - 50;
      if (m1_suit == trump && rank < m1_rank) return -32 + r_rank + (cnt << 6) / 40;
      const int add = (cnt << 6) / 50;
      if (f.high1 == 0) {
        if (max4th != 0) return (follow_top2(f, ls, 1) == lead ? 36 : 48) - rank + add;
        if (bit(card) > (hR & T)) return 48 - rank + add;
        return -12 - rank + add;
      }
      if (max4th != 0) return 72 - rank + add;
      if (bit(card) > (hR & T)) return 48 - rank + add;
      return 36 - rank + add;
    }
    int add = (cnt << 6) / 24;
    if (cnt == 2 && sec_mine) add -= 4;
    if (cnt == 1 && win_mine) add -= 4;
    return -rank + add;
  }

  // Fourth hand.
  const int m2_rank = rank_of(f.move2), m2_suit = f.move2 >> 4;
  if (!is_void) {
    if (f.high2 == 1 || (ls != trump && m2_suit == trump)) return -rank;
    return rank > m2_rank ? 30 - rank : -rank;
  }
  if (trump >= 0) {
    int val = (cnt << 6) / 24;
    if (cnt == 2 && sec_mine) val -= 2;
    if (ls == trump) return -rank + val;
    if (f.high2 == 1) return s == trump ? 2 - rank + val : 25 - rank + val;
    if (m2_suit == trump) {
      if (s == trump) return rank > m2_rank ? 33 + r_rank : -13 + r_rank;
      return 14 - rank + val;
    }
    if (s == trump) return 33 + r_rank;
    return 14 - rank + val;
  }
  int val = (cnt << 6) / 27;
  if (cnt == 2 && sec_mine) val -= 6;
  else if (cnt == 1 && win_mine) val -= 8;
  return -rank + val;
}

}  // namespace ddx
