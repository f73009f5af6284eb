// sr_evogp.h -- conversion between the evogp tree layout (defs.h:30-107)
// and prefix token streams, one individual at a time (host and device).
//
// evogp: per individual, max_len slots of value (float: constant, variable
// index, or the opcode as a float), node type (int16: VAR 0, CONST 1,
// UFUNC 2, BFUNC 3, TFUNC 4; bit 7 marks an output node and is ignored) and
// subtree size (int16, size[0] = tree length), in prefix order.
#pragma once

#include "sr_format.h"

namespace sr {

constexpr int EVOGP_VAR = 0, EVOGP_CONST = 1, EVOGP_UFUNC = 2, EVOGP_BFUNC = 3;
constexpr int EVOGP_TYPE_MASK = 0x7f;

// gpu_simple_jinha_with_evolve.cu:846-866 restated: constants and variables
// keep their value, a function node's value is its opcode.  Returns the
// length, or 0 when size[0] is outside 1..min(max_len, cap).
SR_HD int evogp_to_prefix_one(const float *value, const int16_t *type, const int16_t *size, int max_len,
                              int cap, int32_t *tok, float *val)
{
    const int n = size[0];
    if (n <= 0 || n > max_len || n > cap) return 0;
    for (int t = 0; t < n; ++t) {
        const int nt = type[t] & EVOGP_TYPE_MASK;
        if (nt == EVOGP_CONST) {
            tok[t] = TOK_CONST;
            val[t] = value[t];
        } else if (nt == EVOGP_VAR) {
            tok[t] = TOK_VAR;
            val[t] = value[t];
        } else {
            tok[t] = (int32_t)value[t];
            val[t] = 0.0f;
        }
    }
    return n;
}

// The inverse: node types from token arity, subtree sizes by a reverse scan.
// Returns n, or 0 (size[0] = 0) when n is outside 1..max_len, a token is not
// a constant, variable or operator, or the stream is not exactly one tree.
SR_HD int prefix_to_evogp_one(const int32_t *tok, const float *val, int n, int max_len, float *value,
                              int16_t *type, int16_t *size)
{
    if (n <= 0 || n > max_len || n > 32767) {
        if (max_len > 0) size[0] = 0;
        return 0;
    }
    for (int t = n - 1; t >= 0; --t) {
        const int k = tok[t];
        int s = 1;
        if (k == OP_REF || !is_valid_token(k)) {
            size[0] = 0;
            return 0;
        }
        if (k == TOK_CONST || k == TOK_VAR) {
            type[t] = (int16_t)(k == TOK_CONST ? EVOGP_CONST : EVOGP_VAR);
            value[t] = val[t];
        } else {
            const bool bin = is_binary(k);
            type[t] = (int16_t)(bin ? EVOGP_BFUNC : EVOGP_UFUNC);
            value[t] = (float)k;
            int c = t + 1;
            for (int a = 0; a < (bin ? 2 : 1); ++a) {
                if (c >= n) {
                    size[0] = 0;
                    return 0;
                }
                s += size[c];
                c += size[c];
            }
        }
        size[t] = (int16_t)s;
    }
    if (size[0] != n) {
        size[0] = 0;
        return 0;
    }
    return n;
}

}  // namespace sr
