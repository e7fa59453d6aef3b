// Test support: evaluates the ddx core's key, bound and move-ordering helpers on the
// positions of tests/golden/funcs.npz, filling the same record layout as the fixture
// (tests compile this with g++ and compare field by field).
#include <stdint.h>
#include <string.h>

#include "ddx/search.hpp"

using namespace ddx;

enum { FIELDS = 54 };

extern "C" void ddx_test_funcs(const int64_t* in, int64_t n, int64_t* out) {
  for (int64_t r = 0; r < n; r++) {
    const int64_t* a = in + r * FIELDS;
    int64_t* o = out + r * FIELDS;
    memcpy(o, a, 14 * sizeof(int64_t));  // the position itself
    for (int i = 14; i < FIELDS; i++) o[i] = 0;
    Pos p;
    for (int h = 0; h < 4; h++) p.hand[h] = (u64)a[h];
    p.trump = (int)a[4];
    p.leader = (int)a[5];
    p.nplayed = (int)a[6];
    p.trick = 0;
    for (int i = 0; i < p.nplayed; i++) p.trick |= (u32)a[7 + i] << (8 * i);
    p.ns_won = (int)a[10];
    p.left = (int)a[11];
    const int target = (int)a[12], hint = (int)a[13], need = target - p.ns_won;
    const int seat = (p.leader + p.nplayed) & 3;
    Top t;
    memset(&t, 0, sizeof t);
    if (p.nplayed == 0) {
      u64 code[4];
      suit_codes(p.hand, code);
      const u64 lens = lengths(p.hand);
      for (int s = 0; s < 4; s++) o[14 + s] = (int64_t)code[s];
      o[18] = (int64_t)lens;
      o[19] = (int64_t)key_hash(lens, p.leader);
      top_init(t, p.hand, lens);
      const bool ns_lead = (p.leader & 1) == 0;
      const int cut_ns = need, cut_ew = p.left - need + 1;
      u64 wr = 0;
      bool decided;
      o[20] = quick_tricks(p.hand, t, p.trump, p.leader, ns_lead ? cut_ns : cut_ew,
                           ns_lead ? cut_ew : cut_ns, wr, decided);
      o[21] = (int64_t)wr;
      o[22] = decided;
      wr = 0;
      o[23] = ns_lead ? !later_min(t, p.trump, p.leader, p.ns_won, p.left, target, wr)
                      : later_max(t, p.trump, p.leader, p.ns_won, p.left, target, wr);
      o[24] = (int64_t)wr;
    } else if (p.nplayed == 1) {
      u64 th[4] = {p.hand[0], p.hand[1], p.hand[2], p.hand[3]};
      const int c0 = trick_card(p.trick, 0);
      th[p.leader] |= bit(c0);
      Top t2;
      top_init(t2, th, lengths(th));
      const int second = (p.leader + 1) & 3;
      const int cutoff = (second & 1) == 0 ? need : p.left - need + 1;
      u64 wr = 0;
      o[25] = quick_tricks_2nd(p.hand, t2, p.trump, second, c0, cutoff, wr);
      o[26] = (int64_t)wr;
    }
    Frame f;
    gen_moves(p, seat, hint, t, f);
    o[27] = f.n;
    const u64 own = hand_of(p, seat), all = all_cards(p) | trick_cards(p);
    Follow fw;
    if (p.nplayed) follow_init(fw, p.hand, p.trump, p.leader, p.trick, p.nplayed);
    for (int i = 0; i < 13; i++) {
      o[28 + i] = i < f.n ? move_at(f, i) : -1;
      if (i >= f.n) continue;
      const int card = move_at(f, i);
      const bool hinted = hint >= 0 && equivalent(own, all, hint, card);
      o[41 + i] = p.nplayed == 0 ? lead_weight(p.hand, t, p.trump, p.left, seat, card, hinted)
                                 : hinted ? 1000
                                          : follow_weight(fw, p.hand, p.trump, p.nplayed, seat, card);
    }
  }
}
