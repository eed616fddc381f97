// On-disk formats of nn-merge: data matrices (text, .fvecs, .bvecs), k-NN
// graph text and DPG / layered search graph text.
//
// Readers follow IODelegator (src/iodelegator.cpp):
//   loadMatrix   :186-220  text "row col" then row*col values
//   load_fvecs   :222-265  records [int32 d][d x float32]
//   loadBvec10m  :319-385  records [int32 d][d x uint8] as floats, at most
//                          10,000,000 rows, zero-length records skipped
//   loadKNNGraph :65-102   "N K" then per row "i k id..." (first topk kept)
//   loadDPGraph  :1173-1217 "N" then per row "i deg id..."
// and the layered file of LySearch::loadGraphs (src/lysearch.cpp:112-186):
// DPG blocks one after another, upper layer first.  Writers produce what
// AbstractMerge::saveKGraph (src/abstractmerge.cpp:278-291) and
// LyNNDescent::saveDPGraph (src/lynndescent.cpp:625-638) write.
//
// Every reader is two-phase: a shape call, then a read into caller memory.
// Errors: NNM_ERR_INVALID for unreadable or malformed files (the reference
// prints and exit(0)s, src/iodelegator.cpp:196-201).

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/nnmerge.h"

namespace {

enum Kind { TEXT, FVECS, BVECS, UNKNOWN };

Kind kind_of(const char* path) {
    const std::string p(path);
    auto ends = [&](const char* s) {
        const size_t n = std::strlen(s);
        return p.size() >= n && p.compare(p.size() - n, n, s) == 0;
    };
    if (ends(".fvecs")) return FVECS;
    if (ends(".bvecs")) return BVECS;
    if (ends(".txt")) return TEXT;
    return UNKNOWN;
}

constexpr uint64_t BVECS_MAX_ROWS = 10000000ull;  // loadBvec10m's cap

// Shape of a vecs file: the first record's dimension and the record count
// (file size over record size, as the reference computes it).
int vecs_shape(const char* path, uint32_t elem, uint64_t* rows, uint32_t* dim) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return NNM_ERR_INVALID;
    uint32_t d = 0;
    if (!in.read(reinterpret_cast<char*>(&d), 4) || d == 0) return NNM_ERR_INVALID;
    in.seekg(0, std::ios::end);
    const uint64_t size = (uint64_t)in.tellg();
    *dim = d;
    *rows = size / (4ull + (uint64_t)d * elem);
    return NNM_OK;
}

}  // namespace

extern "C" {

int nnm_io_matrix_shape(const char* path, uint64_t* rows, uint32_t* dim) {
    if (!path || !rows || !dim) return NNM_ERR_INVALID;
    *rows = 0;
    *dim = 0;
    switch (kind_of(path)) {
        case FVECS:
            return vecs_shape(path, 4, rows, dim);
        case BVECS: {
            const int rc = vecs_shape(path, 1, rows, dim);
            if (rc == NNM_OK && *rows > BVECS_MAX_ROWS) *rows = BVECS_MAX_ROWS;
            return rc;
        }
        case TEXT: {
            std::ifstream in(path);
            double r = 0, c = 0;
            if (!(in >> r >> c) || r < 0 || c < 1) return NNM_ERR_INVALID;
            *rows = (uint64_t)(r + 0.5);
            *dim = (uint32_t)(c + 0.5);
            return NNM_OK;
        }
        default:
            return NNM_ERR_INVALID;
    }
}

int nnm_io_read_matrix(const char* path, float* out, uint64_t rows, uint32_t dim) {
    if (!path || (!out && rows)) return NNM_ERR_INVALID;
    uint64_t r0 = 0;
    uint32_t d0 = 0;
    const int rc = nnm_io_matrix_shape(path, &r0, &d0);
    if (rc != NNM_OK) return rc;
    if (d0 != dim || rows > r0) return NNM_ERR_INVALID;
    const Kind k = kind_of(path);
    if (k == TEXT) {
        std::ifstream in(path);
        double hr, hc;
        in >> hr >> hc;
        for (uint64_t i = 0; i < rows * dim; ++i)
            if (!(in >> out[i])) return NNM_ERR_INVALID;
        return NNM_OK;
    }
    std::ifstream in(path, std::ios::binary);
    if (!in) return NNM_ERR_INVALID;
    std::vector<uint8_t> bytes(dim);
    uint64_t got = 0;
    while (got < rows) {
        uint32_t di = 0;
        if (!in.read(reinterpret_cast<char*>(&di), 4)) return NNM_ERR_INVALID;
        if (k == FVECS) {
            if (di != dim) return NNM_ERR_INVALID;  // the reference asserts this
            if (!in.read(reinterpret_cast<char*>(out + got * dim), (std::streamsize)dim * 4)) return NNM_ERR_INVALID;
        } else {
            if (di == 0) continue;  // empty record, skipped (src/iodelegator.cpp:358-359)
            if (di > dim) return NNM_ERR_INVALID;
            if (!in.read(reinterpret_cast<char*>(bytes.data()), di)) return NNM_ERR_INVALID;
            float* row = out + got * dim;
            for (uint32_t j = 0; j < dim; ++j) row[j] = j < di ? (float)bytes[j] : 0.f;
        }
        ++got;
    }
    return NNM_OK;
}

int nnm_io_write_matrix(const char* path, const float* X, uint64_t rows, uint32_t dim) {
    if (!path || (!X && rows) || dim == 0) return NNM_ERR_INVALID;
    const Kind k = kind_of(path);
    if (k == TEXT) {
        FILE* f = std::fopen(path, "w");
        if (!f) return NNM_ERR_INVALID;
        std::fprintf(f, "%llu %u\n", (unsigned long long)rows, dim);
        for (uint64_t i = 0; i < rows; ++i) {
            for (uint32_t j = 0; j < dim; ++j) std::fprintf(f, j ? " %.9g" : "%.9g", X[i * dim + j]);
            std::fputc('\n', f);
        }
        return std::fclose(f) == 0 ? NNM_OK : NNM_ERR_INVALID;
    }
    if (k != FVECS && k != BVECS) return NNM_ERR_INVALID;
    std::ofstream out(path, std::ios::binary);
    if (!out) return NNM_ERR_INVALID;
    std::vector<uint8_t> bytes(dim);
    for (uint64_t i = 0; i < rows; ++i) {
        out.write(reinterpret_cast<const char*>(&dim), 4);
        if (k == FVECS) {
            out.write(reinterpret_cast<const char*>(X + i * dim), (std::streamsize)dim * 4);
        } else {
            for (uint32_t j = 0; j < dim; ++j) {
                const float v = X[i * dim + j];
                if (!(v >= 0.f && v <= 255.f) || v != (float)(int)v) return NNM_ERR_INVALID;
                bytes[j] = (uint8_t)v;
            }
            out.write(reinterpret_cast<const char*>(bytes.data()), dim);
        }
    }
    return out ? NNM_OK : NNM_ERR_INVALID;
}

// ---------------------------------------------------------------- k-NN graph text

int nnm_io_knn_shape(const char* path, uint64_t* rows, uint32_t* k) {
    if (!path || !rows || !k) return NNM_ERR_INVALID;
    std::ifstream in(path);
    uint64_t r = 0;
    uint32_t kk = 0;
    if (!(in >> r >> kk)) return NNM_ERR_INVALID;
    *rows = r;
    *k = kk;
    return NNM_OK;
}

// ids: rows x topk, the first topk ids of every row (rows with fewer ids are
// padded with 0xFFFFFFFF; the reference's calloc leaves 0 there).
int nnm_io_read_knn(const char* path, uint32_t* ids, uint64_t rows, uint32_t topk) {
    if (!path || (!ids && rows) || topk == 0) return NNM_ERR_INVALID;
    std::ifstream in(path);
    uint64_t r = 0;
    uint32_t kk = 0;
    if (!(in >> r >> kk) || rows > r || topk > kk) return NNM_ERR_INVALID;
    for (uint64_t i = 0; i < rows; ++i) {
        uint64_t idx = 0;
        uint32_t k = 0;
        if (!(in >> idx >> k)) return NNM_ERR_INVALID;
        for (uint32_t j = 0; j < k; ++j) {
            uint32_t v = 0;
            if (!(in >> v)) return NNM_ERR_INVALID;
            if (j < topk) ids[i * topk + j] = v;
        }
        for (uint32_t j = k; j < topk; ++j) ids[i * topk + j] = 0xFFFFFFFFu;
    }
    return NNM_OK;
}

int nnm_io_write_knn(const char* path, const uint32_t* ids, uint64_t rows, uint32_t k) {
    if (!path || (!ids && rows) || k == 0) return NNM_ERR_INVALID;
    FILE* f = std::fopen(path, "w");
    if (!f) return NNM_ERR_INVALID;
    std::fprintf(f, "%llu %u\n", (unsigned long long)rows, k);
    for (uint64_t i = 0; i < rows; ++i) {
        std::fprintf(f, "%llu %u ", (unsigned long long)i, k);
        for (uint32_t j = 0; j < k; ++j) std::fprintf(f, "%u ", ids[i * k + j]);
        std::fputc('\n', f);
    }
    return std::fclose(f) == 0 ? NNM_OK : NNM_ERR_INVALID;
}

// ---------------------------------------------------------------- DPG / layers

// Shape of a file of DPG blocks: the number of blocks (at most max_layers)
// and, per block, its rows and edge count.
int nnm_io_layers_shape(const char* path, uint32_t max_layers, uint32_t* nlayers, uint32_t* rows, uint64_t* edges) {
    if (!path || !nlayers || !rows || !edges) return NNM_ERR_INVALID;
    std::ifstream in(path);
    if (!in) return NNM_ERR_INVALID;
    uint32_t g = 0;
    while (g < max_layers) {
        uint64_t r = 0;
        if (!(in >> r)) break;
        uint64_t e = 0;
        for (uint64_t i = 0; i < r; ++i) {
            uint64_t idx = 0, deg = 0;
            if (!(in >> idx >> deg)) return NNM_ERR_INVALID;
            for (uint64_t j = 0; j < deg; ++j) {
                uint32_t v;
                if (!(in >> v)) return NNM_ERR_INVALID;
            }
            e += deg;
        }
        rows[g] = (uint32_t)r;
        edges[g] = e;
        ++g;
    }
    *nlayers = g;
    return g ? NNM_OK : NNM_ERR_INVALID;
}

// Reads nlayers blocks into CSR: offs[g] (rows[g] + 1 entries), adj[g].
int nnm_io_read_layers(const char* path, uint32_t nlayers, uint32_t* const* offs, uint32_t* const* adj) {
    if (!path || !offs || !adj) return NNM_ERR_INVALID;
    std::ifstream in(path);
    if (!in) return NNM_ERR_INVALID;
    for (uint32_t g = 0; g < nlayers; ++g) {
        uint64_t r = 0;
        if (!(in >> r) || !offs[g]) return NNM_ERR_INVALID;
        uint64_t e = 0;
        offs[g][0] = 0;
        for (uint64_t i = 0; i < r; ++i) {
            uint64_t idx = 0, deg = 0;
            if (!(in >> idx >> deg)) return NNM_ERR_INVALID;
            for (uint64_t j = 0; j < deg; ++j) {
                uint32_t v;
                if (!(in >> v)) return NNM_ERR_INVALID;
                if (adj[g]) adj[g][e + j] = v;
            }
            e += deg;
            offs[g][i + 1] = (uint32_t)e;
        }
    }
    return NNM_OK;
}

// Writes one DPG block ("N" then "i deg ids"), appended when append != 0.
int nnm_io_write_layer(const char* path, int append, const uint32_t* offs, const uint32_t* adj, uint32_t rows) {
    if (!path || !offs || (!adj && offs[rows])) return NNM_ERR_INVALID;
    FILE* f = std::fopen(path, append ? "a" : "w");
    if (!f) return NNM_ERR_INVALID;
    std::fprintf(f, "%u\n", rows);
    for (uint32_t i = 0; i < rows; ++i) {
        std::fprintf(f, "%u %u ", i, offs[i + 1] - offs[i]);
        for (uint32_t e = offs[i]; e < offs[i + 1]; ++e) std::fprintf(f, "%u ", adj[e]);
        std::fputc('\n', f);
    }
    return std::fclose(f) == 0 ? NNM_OK : NNM_ERR_INVALID;
}

}  // extern "C"
