// sr_io.cpp -- the reference's on-disk formats, read and written on the host.
//
// Digest ("gen_k.txt", parse_input_info, utils.cpp:122-233):
//   num_exprs
//   then per expression: num_vars num_dps num_tokens, the tokens, the values
//   (doubles), and the data file path on a line of its own.
// Data file ("shared_data.txt", load_data_file, utils.cpp:235-300): one
//   datapoint per row, num_vars + 1 numbers (x0 .. x_{nv-1}, y), parsed with
//   strtod and returned column by column.
// Reading follows the reference's stream operations (operator>> on an
// ifstream for the digest, strtod for data), so every parsed number is
// bit-identical to the reference's; unlike the reference, a truncated or
// malformed digest is an error instead of zero-filled arrays.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/sr_hip.h"

struct sr_digest {
    std::vector<int32_t> nv, ndps, off, len, tok;
    std::vector<double> val;
    std::vector<std::string> path;
    bool shared = false;
};

extern "C" {

int sr_digest_read(const char *path, sr_digest **out)
{
    if (!path || !out) return SR_ERR_INVALID_ARG;
    *out = nullptr;
    std::ifstream f(path);
    if (!f.is_open()) return SR_ERR_IO;
    auto *d = new sr_digest;
    int E = 0;
    f >> E;
    if (!f || E <= 0) {
        delete d;
        return SR_ERR_BAD_EXPR;
    }
    d->nv.resize(E);
    d->ndps.resize(E);
    d->off.resize(E);
    d->len.resize(E);
    d->path.resize(E);
    for (int e = 0; e < E; ++e) {
        int nv = 0, ndps = 0, n = 0;
        f >> nv >> ndps >> n;
        if (!f || nv < 0 || ndps < 0 || n <= 0 || n > (1 << 20)) {
            delete d;
            return SR_ERR_BAD_EXPR;
        }
        d->nv[e] = nv;
        d->ndps[e] = ndps;
        d->off[e] = (int32_t)d->tok.size();
        d->len[e] = n;
        for (int i = 0; i < n; ++i) {
            int t = 0;
            f >> t;
            d->tok.push_back(t);
        }
        for (int i = 0; i < n; ++i) {
            double v = 0.0;
            f >> v;
            d->val.push_back(v);
        }
        if (!f) {
            delete d;
            return SR_ERR_BAD_EXPR;
        }
        f.ignore();  // the newline after the values, as utils.cpp:196-197
        std::getline(f, d->path[e]);
    }
    // utils.cpp:202-223: shared when there is more than one expression and
    // every one names the same file, variable count and datapoint count
    if (E > 1) {
        bool same = true;
        for (int e = 1; e < E && same; ++e)
            same = d->path[e] == d->path[0] && d->nv[e] == d->nv[0] && d->ndps[e] == d->ndps[0];
        d->shared = same;
    }
    *out = d;
    return SR_OK;
}

int sr_digest_free(sr_digest *d)
{
    delete d;
    return SR_OK;
}

int sr_digest_shape(const sr_digest *d, int32_t *num_exprs, int64_t *num_tokens, int32_t *shared_data)
{
    if (!d) return SR_ERR_INVALID_ARG;
    if (num_exprs) *num_exprs = (int32_t)d->len.size();
    if (num_tokens) *num_tokens = (int64_t)d->tok.size();
    if (shared_data) *shared_data = d->shared ? 1 : 0;
    return SR_OK;
}

int sr_digest_population(const sr_digest *d, int32_t *tokens, double *values, int32_t *offsets,
                         int32_t *lengths, int32_t *num_vars, int32_t *num_dps)
{
    if (!d) return SR_ERR_INVALID_ARG;
    const size_t E = d->len.size(), T = d->tok.size();
    if (tokens) std::memcpy(tokens, d->tok.data(), T * 4);
    if (values) std::memcpy(values, d->val.data(), T * 8);
    if (offsets) std::memcpy(offsets, d->off.data(), E * 4);
    if (lengths) std::memcpy(lengths, d->len.data(), E * 4);
    if (num_vars) std::memcpy(num_vars, d->nv.data(), E * 4);
    if (num_dps) std::memcpy(num_dps, d->ndps.data(), E * 4);
    return SR_OK;
}

int sr_digest_data_path(const sr_digest *d, int32_t expr, char *out, int64_t cap, int64_t *len)
{
    if (!d || expr < 0 || expr >= (int32_t)d->path.size()) return SR_ERR_INVALID_ARG;
    const std::string &p = d->path[expr];
    if (len) *len = (int64_t)p.size();
    if (out && cap > 0) {
        const size_t n = std::min<size_t>(p.size(), (size_t)cap - 1);
        std::memcpy(out, p.data(), n);
        out[n] = '\0';
    }
    return SR_OK;
}

int sr_digest_write(const char *path, const sr_population *pop, const int32_t *num_vars, const int32_t *num_dps,
                    const char *const *data_paths)
{
    if (!path || !pop || !num_vars || !num_dps || !data_paths || pop->num_exprs <= 0 || !pop->tokens ||
        !pop->offsets || !pop->lengths || (!pop->values && !pop->values_f32))
        return SR_ERR_INVALID_ARG;
    FILE *f = std::fopen(path, "w");
    if (!f) return SR_ERR_IO;
    std::fprintf(f, "%d\n", pop->num_exprs);
    for (int e = 0; e < pop->num_exprs; ++e) {
        const int32_t o = pop->offsets[e], n = pop->lengths[e];
        std::fprintf(f, "%d %d %d\n", num_vars[e], num_dps[e], n);
        for (int i = 0; i < n; ++i) std::fprintf(f, i ? " %d" : "%d", pop->tokens[o + i]);
        std::fputc('\n', f);
        for (int i = 0; i < n; ++i) {
            const double v = pop->values ? pop->values[o + i] : (double)pop->values_f32[o + i];
            std::fprintf(f, i ? " %.17g" : "%.17g", v);  // round-trips every double
        }
        std::fprintf(f, "\n%s\n", data_paths[e] ? data_paths[e] : "");
    }
    const bool ok = std::ferror(f) == 0;
    return (std::fclose(f) == 0 && ok) ? SR_OK : SR_ERR_IO;
}

int sr_data_file_read(const char *path, int32_t num_vars, int64_t num_dps, double *out, int64_t *parsed)
{
    if (!path || !out || num_vars < 0 || num_dps < 0) return SR_ERR_INVALID_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) return SR_ERR_IO;
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<char> buf((size_t)std::max(size, 0L) + 1);
    const size_t got = size > 0 ? std::fread(buf.data(), 1, (size_t)size, f) : 0;
    std::fclose(f);
    if ((long)got != size) return SR_ERR_IO;
    buf[got] = '\0';
    // utils.cpp:279-292: strtod per number, whitespace skipped after each
    char *p = buf.data();
    char *const end = buf.data() + got;
    int64_t ok = 0;
    for (int64_t j = 0; j < num_dps; ++j) {
        for (int32_t i = 0; i <= num_vars; ++i) {
            char *q = p;
            out[(int64_t)i * num_dps + j] = std::strtod(p, &q);
            ok += q != p;
            p = q;
            while (p < end && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r')) ++p;
        }
    }
    if (parsed) *parsed = ok;
    return SR_OK;
}

int sr_data_file_write(const char *path, int32_t num_vars, int64_t num_dps, const double *cols)
{
    if (!path || !cols || num_vars < 0 || num_dps < 0) return SR_ERR_INVALID_ARG;
    FILE *f = std::fopen(path, "w");
    if (!f) return SR_ERR_IO;
    for (int64_t j = 0; j < num_dps; ++j) {
        for (int32_t i = 0; i <= num_vars; ++i)
            std::fprintf(f, i ? " %.17g" : "%.17g", cols[(int64_t)i * num_dps + j]);
        std::fputc('\n', f);
    }
    const bool ok = std::ferror(f) == 0;
    return (std::fclose(f) == 0 && ok) ? SR_OK : SR_ERR_IO;
}

}  // extern "C"
