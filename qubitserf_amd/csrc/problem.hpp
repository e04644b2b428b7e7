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
    int qubits = 0;         // > 0: symplectic problem, n = 2 qubits columns [z | x] and the
                            // weight of a vector is the number of qubits it touches
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
// The same problem in symplectic form (cpp/distfind/stab.cpp:98-132): gen = the
// centralizer of S over 2q columns, detectors pre-swapped so that <d, c> is a plain dot
// product, qubits = q, seed_upper = the lightest logical row's symplectic weight.
Problem symplectic_plain(const BitMat& S);
bool is_css_symplectic(const BitMat& S);
void split_css(const BitMat& S, BitMat& Hx, BitMat& Hz);

// Nonzero columns of gen.
int nonzero_columns(const BitMat& gen);

// The Edmonds matroid partition of gen's columns (cpp/distfind/css.cpp:274-328): classes
// in the reference's order (unstable sort by size, css.cpp:354-355), empty classes
// dropped. false when the search fails its consistency check (css.cpp:322-327). The
// information sets themselves are eliminated on the device (qsf_hip_eliminate).
bool matroid_parts(const BitMat& gen, std::vector<std::vector<int>>& parts);

}  // namespace qsa
