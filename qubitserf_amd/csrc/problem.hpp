// Distance subproblems and information sets (host side).
//
// Every distance question the engine answers is reduced to one shape (the reference's
// DistProblem, cpp/distfind/css.hpp:21-30): the minimum Hamming weight of a nonzero vector
// in rowspan(gen) that is a logical (some detector row has odd overlap with it, or every
// nonzero vector counts).
#pragma once
#include <vector>
#include "gf2.hpp"

namespace qsa {

struct Problem {
    int n = 0;
    BitMat gen;             // K x n basis of the space searched
    BitMat detector;        // rows d with <d, c> = 1 for some row  <=>  c is a logical
    bool every_nonzero = false;  // classical: every nonzero codeword is a logical
    bool even = false;      // every codeword has even weight
    int seed_upper = 0;     // a known logical's weight (0 = none)
};

// CSS side 'Z' searches ker(Hx) modulo rowspan(Hz); side 'X' the converse
// (cpp/distfind/css.cpp:78-87, :107-109).
Problem css_side(const BitMat& Hx, const BitMat& Hz, char side);
// Classical code with parity check H (cpp/distfind/css.cpp:111-120).
Problem classical(const BitMat& H);

// Non-CSS stabilizer S (m x 2q, [z|x]): the symplectic problem mapped through
// (a|b) -> (a|b|a^b) to a Hamming problem of length 3q whose distance is twice the
// symplectic one (cpp/distfind/stab.cpp:98-132, :147-178).
Problem symplectic_isometric(const BitMat& S);
bool is_css_symplectic(const BitMat& S);
void split_css(const BitMat& S, BitMat& Hx, BitMat& Hz);

// One information set: a generator of the same code that is the identity on the set's
// pivot columns (rows [0, rank)) and zero there on rows [rank, K).
struct InfoSet {
    BitMat gen;
    int rank = 0;
    std::vector<int> pivots;   // pivot column of row i, i < rank
};

// Disjoint information sets covering the nonzero columns: greedy peeling (each round
// eliminates over the columns left over by the previous rounds), replaced by an
// Edmonds matroid partition when n <= 200 and the partition yields more full-rank sets
// (cpp/distfind/css.cpp:177-203, :274-328, :336-366). Returned sorted by rank,
// descending, stable.
std::vector<InfoSet> information_sets(const BitMat& gen);
// Same, from an explicit column partition (each part eliminated over its own columns).
std::vector<InfoSet> information_sets_from_parts(const BitMat& gen, const std::vector<std::vector<int>>& parts);

}  // namespace qsa
