// pbn: one deal of a Pgx .npy file as a PBN deal string, followed by its stored DD table.
//   pbn FILE.npy I
// Line 1 is a PBN "Deal" value: "N:" then the four hands clockwise from North, separated by
// spaces; a hand is its spades, hearts, diamonds and clubs joined by '.', each suit's cards
// written high to low with the PBN rank letters A K Q J T 9 .. 2. Line 2 is the stored table
// as 20 comma-separated trick counts, declarers N, E, S, W in turn and, for each, the strains
// C, D, H, S, NT (the reference's tests/pbn.cpp prints the same two lines). Exit status: 0 on
// success, 1 when the file or the deal cannot be read, 2 on a usage error.
#include <stdio.h>
#include <stdlib.h>

#include <string>

#include "ddx.h"

namespace {

// PBN rank letter of rank index r (0 = deuce .. 12 = ace).
char rank_letter(int r) { return r < 8 ? (char)('2' + r) : "TJQKA"[r - 8]; }

// "S.H.D.C" text of one hand (bit suit * 16 + rank, suit 0 = spades).
std::string hand_text(uint64_t hand) {
  std::string text;
  for (int suit = 0; suit < 4; ++suit) {
    if (suit > 0) text.push_back('.');
    unsigned ranks = (unsigned)(hand >> (16 * suit)) & 0x1FFFu;
    while (ranks != 0) {  // highest remaining rank first
      const int r = 31 - __builtin_clz(ranks);
      text.push_back(rank_letter(r));
      ranks &= ~(1u << r);
    }
  }
  return text;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s FILE.npy I\n", argv[0]);
    return 2;
  }
  uint64_t hands[4];
  uint8_t table[20];
  int64_t rows = 0;
  if (ddx_pgx_read(argv[1], atoll(argv[2]), 1, hands, table, &rows) != DDX_OK || rows != 1) {
    fprintf(stderr, "pbn: cannot read deal %s of %s\n", argv[2], argv[1]);
    return 1;
  }
  std::string deal = "N:";
  for (int seat = 0; seat < 4; ++seat) deal += (seat > 0 ? " " : "") + hand_text(hands[seat]);
  std::string tricks;
  for (int declarer = 0; declarer < 4; ++declarer)
    for (int strain = 0; strain < 5; ++strain) {
      if (!tricks.empty()) tricks.push_back(',');
      tricks += std::to_string(table[4 * strain + declarer]);
    }
  printf("%s\n%s\n", deal.c_str(), tricks.c_str());
  return 0;
}
