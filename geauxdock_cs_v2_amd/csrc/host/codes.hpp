// Type code tables of the GeauxDock force field.
//
// The numbering must match the parameter files, so the tables keep the
// reference's code assignment (src/frontend/data.cpp:238-587): residue codes
// by three- and one-letter name, effective-point codes, and the 24 ligand
// SYBYL atom types.  Unknown ligand types map to GD_BADKDE like the
// reference's getLigCode; unknown residues or points are an error.
#pragma once

#include <string>

namespace gd {

// Returns 0..19, or -1 for an unknown residue name.
inline int residue_code3(const std::string &s)
{
    static const char *names[20] = {"ALA", "ARG", "CYS", "PRO", "LEU", "ILE", "GLU",
                                    "HIS", "TRP", "THR", "MET", "ASP", "GLY", "PHE",
                                    "LYS", "GLN", "SER", "ASN", "TYR", "VAL"};
    for (int i = 0; i < 20; ++i)
        if (s == names[i]) return i;
    return -1;
}

inline int residue_code1(const std::string &s)
{
    static const char *names = "ARCPLIEHWTMDGFKQSNYV";
    if (s.size() != 1) return -1;
    for (int i = 0; i < 20; ++i)
        if (s[0] == names[i]) return i;
    return -1;
}

inline const char *residue_name3(int code)
{
    static const char *names[20] = {"ALA", "ARG", "CYS", "PRO", "LEU", "ILE", "GLU",
                                    "HIS", "TRP", "THR", "MET", "ASP", "GLY", "PHE",
                                    "LYS", "GLN", "SER", "ASN", "TYR", "VAL"};
    return (code >= 0 && code < 20) ? names[code] : "UNK";
}

// Effective point codes 0..29 (CA, PP, side-chain points), -1 if unknown.
inline int point_code(const std::string &s)
{
    static const char *names[30] = {"CA", "PP", "A1", "C1", "D1", "E1", "E2", "F1", "F2", "H1",
                                    "H2", "I1", "K1", "K2", "L1", "M1", "M2", "N1", "P1", "Q1",
                                    "Q2", "R1", "R2", "S1", "T1", "V1", "W1", "W2", "Y1", "Y2"};
    for (int i = 0; i < 30; ++i)
        if (s == names[i]) return i;
    return -1;
}

// Ligand SYBYL atom type codes 0..23; GD_BADKDE (50) for anything else.
inline int ligand_code(const std::string &s)
{
    static const char *names[24] = {"Br",  "C.1",  "C.2", "C.3",  "C.ar", "C.cat",
                                    "Cl",  "F",    "I",   "N.1",  "N.2",  "N.3",
                                    "N.4", "N.am", "N.ar", "N.pl3", "O.2", "O.3",
                                    "O.co2", "P.3", "S.2", "S.3",  "S.O",  "S.O2"};
    for (int i = 0; i < 24; ++i)
        if (s == names[i]) return i;
    return 50;
}

inline const char *ligand_name(int code)
{
    static const char *names[24] = {"Br",  "C.1",  "C.2", "C.3",  "C.ar", "C.cat",
                                    "Cl",  "F",    "I",   "N.1",  "N.2",  "N.3",
                                    "N.4", "N.am", "N.ar", "N.pl3", "O.2", "O.3",
                                    "O.co2", "P.3", "S.2", "S.3",  "S.O",  "S.O2"};
    return (code >= 0 && code < 24) ? names[code] : "Du";
}

// Side-chain effective points generated for a residue code: up to two
// (point type, class) pairs, class 2 = first side-chain point, class 3 =
// second (src/frontend/load.cpp:631-698).  GLY (12) has none.
struct SideChain {
    int n;
    int type[2];
    int cls[2];
};

inline SideChain side_chain_points(int res)
{
    switch (res) {
    case 0: return {1, {2, 0}, {2, 0}};
    case 1: return {2, {21, 22}, {2, 3}};
    case 2: return {1, {3, 0}, {2, 0}};
    case 3: return {1, {18, 0}, {2, 0}};
    case 4: return {1, {14, 0}, {2, 0}};
    case 5: return {1, {11, 0}, {2, 0}};
    case 6: return {2, {5, 6}, {2, 3}};
    case 7: return {2, {9, 10}, {2, 3}};
    case 8: return {2, {26, 27}, {2, 3}};
    case 9: return {1, {24, 0}, {2, 0}};
    case 10: return {2, {15, 16}, {2, 3}};
    case 11: return {1, {4, 0}, {2, 0}};
    case 13: return {2, {7, 8}, {2, 3}};
    case 14: return {2, {12, 13}, {2, 3}};
    case 15: return {2, {19, 20}, {2, 3}};
    case 16: return {1, {23, 0}, {2, 0}};
    case 17: return {1, {17, 0}, {2, 0}};
    case 18: return {2, {28, 29}, {2, 3}};
    case 19: return {1, {25, 0}, {2, 0}};
    default: return {0, {0, 0}, {0, 0}};
    }
}

// Which PDB atoms average into an effective point of a given point type
// (src/frontend/load.cpp:760-805): CA point = the CA atom; single side-chain
// points = every atom except backbone N/CA/C/O; first point of a two-point
// side chain = every atom except N/CA/C/O/CB/CG; second point = CB and CG.
enum PointAtomRule { RULE_CA, RULE_SC_ALL, RULE_SC_DISTAL, RULE_SC_PROXIMAL, RULE_PP };

inline PointAtomRule point_atom_rule(int t)
{
    switch (t) {
    case 0: return RULE_CA;
    case 1: return RULE_PP;
    case 2: case 3: case 4: case 11: case 14: case 17: case 18: case 23: case 24: case 25:
        return RULE_SC_ALL;
    case 5: case 7: case 9: case 12: case 15: case 19: case 21: case 26: case 28:
        return RULE_SC_DISTAL;
    default:
        return RULE_SC_PROXIMAL;
    }
}

} // namespace gd
