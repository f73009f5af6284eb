// sr_adjoint_prog.h -- host-side preparation of parametric programs for the
// reverse (adjoint) pass.
//
// A parametric program is an accumulator program (compile_expr) whose
// constant operands may name a free-constant slot (param_postpass).  Value
// v_i = the accumulator after instruction i; every value is consumed exactly
// once: as the previous top of instruction i + 1 (a binary op, or a unary op
// on the top), or as the operand a later SRC_POP instruction pops.
//
// adjoint_postpass marks each instruction whose value depends on a free
// constant (active), assigns compact tape rows to the values the reverse pass
// reads (the output and the operands of every active instruction), rewrites a
// SRC_POP's arg to the tape row of the popped value, and fills aux[i]:
//   active i   : tape row of v_{i-1} when it is an operand of i, else -1;
//   inactive i : first instruction of i's subtree -- the reverse pass jumps
//                over the whole range (a subtree's instructions are
//                contiguous and it has no free constant).
#pragma once

#include <vector>

#include "sr_format.h"

namespace sr {

constexpr int BC_ACTIVE = 1 << 15;   // value depends on a free constant
constexpr int BC_TAPE_SHIFT = 20;    // bits 20..30: 1 + tape row of the value (0: not taped)
constexpr int kMaxTapeRows = 0x7ff;
SR_HD int bc_tape(int code) { return ((code >> BC_TAPE_SHIFT) & 0x7ff) - 1; }

// Uses the previous top as an operand: binary ops and unary ops on the top.
SR_HD bool uses_prev(int code)
{
    const int cls = bc_cls(code);
    return cls == CL_BIN || (cls == CL_UN && bc_src(code) == SRC_NONE);
}

// Returns the number of tape rows, or -1 when more than kMaxTapeRows are needed.
inline int adjoint_postpass(Ins *ins, int n, int32_t *aux)
{
    std::vector<char> active(n, 0), taped(n, 0);
    std::vector<int> start(n, 0), row(n, -1);
    for (int i = 0; i < n; ++i) {
        const int code = ins[i].code;
        const int src = bc_src(code);
        bool act = src == SRC_CONST && bc_slot(code) >= 0;
        int s = i;
        if (uses_prev(code)) {
            act = act || active[i - 1];
            s = start[i - 1];
        }
        if (src == SRC_POP) {
            const int p = ins[i].arg;  // producer index (param_postpass)
            act = act || active[p];
            s = start[p];
        }
        active[i] = act;
        start[i] = s;
    }
    for (int i = 0; i < n; ++i) {
        if (!active[i]) continue;
        taped[i] = 1;
        if (uses_prev(ins[i].code)) taped[i - 1] = 1;
        if (bc_src(ins[i].code) == SRC_POP) taped[ins[i].arg] = 1;
    }
    int rows = 0;
    for (int i = 0; i < n; ++i)
        if (taped[i]) row[i] = rows++;
    if (rows > kMaxTapeRows) return -1;
    for (int i = 0; i < n; ++i) {
        int code = ins[i].code & ~(BC_ACTIVE | (0x7ff << BC_TAPE_SHIFT));
        if (active[i]) code |= BC_ACTIVE;
        if (taped[i]) code |= (row[i] + 1) << BC_TAPE_SHIFT;
        if (active[i]) aux[i] = uses_prev(code) ? row[i - 1] : -1;
        else aux[i] = start[i];
        if (bc_src(code) == SRC_POP) ins[i].arg = row[ins[i].arg];  // -1 when the popped value is not taped
        ins[i].code = code;
    }
    return rows;
}

}  // namespace sr
