// code.h -- runtime parity-check code description and the level planner.
//
// The reference fixes one code per binary at compile time
// (src/matrix/code.h:1, src/matrix/constantes_gpu.h:17-75): a table
// PosNoeudsVariable[_M] listing, check by check in processing order, the
// variables of each check, with at most two check degrees.  Here the same
// table is a runtime object with any degree per check, so one process can
// hold many codes.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

struct ldpc_code {
    int N = 0;  // variables
    int C = 0;  // checks
    int E = 0;  // edges
    std::vector<int> deg;         // [C] table order
    std::vector<uint32_t> pos;    // [E] variable of each edge, table order
    std::vector<size_t> eoff;     // [C+1] first edge of each check
    int q1_split = 0;             // first check of the second degree group
    int max_deg = 0;
};

namespace ldpc {

// Validate and finish a code (eoff, max_deg, q1_split).  Returns "" or an error.
std::string finish_code(ldpc_code &c, int q1_split);

// Dependency levels of the serial table order: level(c) = 1 + max level of
// the earlier checks sharing a variable with c.  Checks of one level share no
// variable, and processing levels in order reproduces the serial order.
int dependency_levels(const ldpc_code &c, std::vector<int> *level_of);

// Level-sorted execution plan of a code.
struct Plan {
    int n_levels = 0;
    int max_deg = 0;
    std::vector<int> level_beg;         // [n_levels + 1]
    std::vector<int> level_deg;         // [n_levels] the degree every check of the level has, or 0
    std::vector<int> order;             // level-sorted -> table check index
    std::vector<int> slot_of;           // table check index -> level-sorted
    std::vector<uint32_t> check_info;   // [C] degree | q1 << 8
    std::vector<uint32_t> check_e0;     // [C] first level-order edge
    std::vector<uint32_t> edge_var;     // [E] level-order
};
// granule <= 1: the levels of dependency_levels.  granule > 1: the same number
// of levels, with checks that have slack moved so that level sizes are
// multiples of `granule` where the dependencies allow (a kernel that maps
// (check, word) items of a level onto 64-lane waves passes 64 / words per
// block, so that the waves at a level's end are full); decisions are the same.
Plan make_plan(const ldpc_code &c, int granule = 1);

// generators (deterministic)
std::string gen_regular(ldpc_code &out, int n, int dv, int dc, uint64_t seed);
std::string gen_qc(ldpc_code &out, int mb, int nb, int z, double p, uint64_t seed);
std::string gen_ira(ldpc_code &out, int k, int n_hi, int d_hi, int d_lo, uint64_t seed);
std::string reorder_by_colour(const ldpc_code &in, ldpc_code &out);
std::string load_table(ldpc_code &out, const char *path, int n_vars);

}  // namespace ldpc
