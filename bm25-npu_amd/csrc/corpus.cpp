// Synthetic corpus generator and BM25 index-time weights (host C++).
//
// The generator implements this project's seeded corpus spec: token ranks
// drawn from Zipf(s) over the vocabulary, document lengths from a lognormal of
// given mean and sigma, tf = multiplicity of a term in its document. The
// reference's own generator (standalone/large_corpus.py:58-62) draws uniform
// tokens from Python's Mersenne Twister; its index layout, the CSC arrays
// term_offsets/doc_ids/tf/doc_norm/idf (standalone/resident_index.py:12-18),
// and its weight formulas (large_corpus.py:148,174) are what this file keeps.
//
// Every document has its own counter-seeded random stream, so the output is
// a function of the spec alone, independent of the thread count.
#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>

#include "common.h"

namespace bm25mi {

ErrorInfo &ThreadError()
{
    static thread_local ErrorInfo info;
    return info;
}

namespace {

inline uint64_t SplitMix(uint64_t &x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// xoshiro256** seeded from (seed, stream) through SplitMix64.
struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed ^ (stream * 0xd1342543de82ef95ULL + 0x2545f4914f6cdd1dULL);
        for (auto &v : s)
            v = SplitMix(x);
    }
    static inline uint64_t Rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    inline uint64_t Next()
    {
        const uint64_t result = Rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = Rotl(s[3], 45);
        return result;
    }
    // Uniform in [0, 1) with 53 random bits.
    inline double Uniform() { return static_cast<double>(Next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Vose alias table over ranks 1..V with weight r^-s.
struct Alias {
    std::vector<double> prob;
    std::vector<uint32_t> alias;
    Alias(int64_t v, double s)
    {
        prob.resize(v);
        alias.resize(v);
        std::vector<double> w(v);
        double total = 0;
        for (int64_t r = 0; r < v; ++r) {
            w[r] = std::pow(static_cast<double>(r + 1), -s);
            total += w[r];
        }
        std::vector<uint32_t> small, large;
        small.reserve(v);
        large.reserve(v);
        for (int64_t r = 0; r < v; ++r) {
            w[r] = w[r] * static_cast<double>(v) / total;
            (w[r] < 1.0 ? small : large).push_back(static_cast<uint32_t>(r));
        }
        while (!small.empty() && !large.empty()) {
            const uint32_t l = small.back(), g = large.back();
            small.pop_back();
            prob[l] = w[l];
            alias[l] = g;
            w[g] = (w[g] + w[l]) - 1.0;
            if (w[g] < 1.0) {
                large.pop_back();
                small.push_back(g);
            }
        }
        for (uint32_t g : large) {
            prob[g] = 1.0;
            alias[g] = g;
        }
        for (uint32_t l : small) {
            prob[l] = 1.0;
            alias[l] = l;
        }
    }
    inline uint32_t Draw(Rng &rng) const
    {
        const double u = rng.Uniform() * static_cast<double>(prob.size());
        const uint64_t i = static_cast<uint64_t>(u);
        const uint64_t col = i < prob.size() ? i : prob.size() - 1;
        const double frac = u - static_cast<double>(col);
        return frac < prob[col] ? static_cast<uint32_t>(col) : alias[col];
    }
};

} // namespace

struct Corpus {
    int64_t docs = 0, vocab = 0, postings = 0;
    std::vector<int64_t> offsets;
    std::vector<int32_t> ids;
    std::vector<float> tf;
    std::vector<int32_t> len;
};

static Corpus *Generate(const bm25mi_corpus_spec &spec)
{
    Require(spec.num_docs >= 1 && spec.num_docs <= INT32_MAX, "num_docs must be in [1, INT32_MAX]");
    Require(spec.vocab >= 1 && spec.vocab < INT32_MAX, "vocab must be in [1, INT32_MAX)");
    Require(std::isfinite(spec.zipf_s) && spec.zipf_s > 0, "zipf_s must be finite and > 0");
    Require(std::isfinite(spec.mean_len) && spec.mean_len >= 1 && spec.mean_len <= 1e6, "mean_len must be in [1, 1e6]");
    Require(std::isfinite(spec.sigma) && spec.sigma >= 0 && spec.sigma <= 4, "sigma must be in [0, 4]");
    const int threads = Threads(spec.threads);
    const int64_t n = spec.num_docs, v = spec.vocab;
    const double mu = std::log(spec.mean_len) - 0.5 * spec.sigma * spec.sigma;
    const Alias zipf(v, spec.zipf_s);
    auto corpus = std::make_unique<Corpus>();
    corpus->docs = n;
    corpus->vocab = v;
    corpus->len.assign(n, 0);

    // Pass 1: per chunk, each document's sorted unique (term, tf) pairs and per-term counts.
    struct Chunk {
        std::vector<uint32_t> terms;
        std::vector<uint32_t> tfs;
        std::vector<uint32_t> perDoc;
        std::vector<uint32_t> df;
    };
    std::vector<Chunk> chunks(threads);
    ParallelChunks(n, threads, [&](int c, int64_t b, int64_t e) {
        Chunk &ch = chunks[c];
        ch.df.assign(v, 0);
        ch.perDoc.reserve(e - b);
        ch.terms.reserve(static_cast<size_t>((e - b) * spec.mean_len));
        ch.tfs.reserve(static_cast<size_t>((e - b) * spec.mean_len));
        std::vector<uint32_t> tokens;
        for (int64_t d = b; d < e; ++d) {
            Rng rng(spec.seed, static_cast<uint64_t>(d));
            // Box-Muller normal from two uniforms; u1 in (0, 1].
            const double u1 = 1.0 - rng.Uniform(), u2 = rng.Uniform();
            const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            double lf = std::exp(mu + spec.sigma * z);
            if (!(lf < 1e7))
                lf = 1e7;
            int64_t length = std::llround(lf);
            if (length < 1)
                length = 1;
            corpus->len[d] = static_cast<int32_t>(length);
            tokens.resize(length);
            for (auto &t : tokens)
                t = zipf.Draw(rng);
            std::sort(tokens.begin(), tokens.end());
            uint32_t unique = 0;
            for (size_t i = 0; i < tokens.size();) {
                size_t j = i;
                while (j < tokens.size() && tokens[j] == tokens[i])
                    ++j;
                ch.terms.push_back(tokens[i]);
                ch.tfs.push_back(static_cast<uint32_t>(j - i));
                ++ch.df[tokens[i]];
                ++unique;
                i = j;
            }
            ch.perDoc.push_back(unique);
        }
    });

    // Term offsets and each chunk's write cursor per term (chunks in doc order keep IDs sorted).
    corpus->offsets.assign(v + 1, 0);
    for (int64_t t = 0; t < v; ++t) {
        int64_t df = 0;
        for (auto &ch : chunks)
            df += ch.df[t];
        corpus->offsets[t + 1] = corpus->offsets[t] + df;
    }
    const int64_t m = corpus->offsets[v];
    Require(m >= 1 && m <= INT32_MAX, "Posting count must be in [1, INT32_MAX]");
    corpus->postings = m;
    ParallelChunks(v, threads, [&](int, int64_t b, int64_t e) {
        for (int64_t t = b; t < e; ++t) {
            int64_t cursor = corpus->offsets[t];
            for (auto &ch : chunks) {
                const uint32_t count = ch.df[t];
                ch.df[t] = static_cast<uint32_t>(cursor - corpus->offsets[t]); // relative start
                cursor += count;
            }
        }
    });
    corpus->ids.resize(m);
    corpus->tf.resize(m);
    ParallelChunks(n, threads, [&](int c, int64_t b, int64_t e) {
        Chunk &ch = chunks[c];
        size_t k = 0;
        for (int64_t d = b; d < e; ++d) {
            const uint32_t unique = ch.perDoc[d - b];
            for (uint32_t u = 0; u < unique; ++u, ++k) {
                const uint32_t t = ch.terms[k];
                const int64_t pos = corpus->offsets[t] + ch.df[t]++;
                corpus->ids[pos] = static_cast<int32_t>(d);
                corpus->tf[pos] = static_cast<float>(ch.tfs[k]);
            }
        }
        std::vector<uint32_t>().swap(ch.terms);
        std::vector<uint32_t>().swap(ch.tfs);
    });
    return corpus.release();
}

} // namespace bm25mi

using namespace bm25mi;

struct bm25mi_corpus : Corpus {
};

extern "C" int bm25mi_corpus_generate(const bm25mi_corpus_spec *spec, bm25mi_corpus **out)
{
    return Invoke([&] {
        Require(spec && out, "Null corpus spec/output");
        *out = nullptr;
        Corpus *c = Generate(*spec);
        auto *wrapped = new bm25mi_corpus();
        static_cast<Corpus &>(*wrapped) = std::move(*c);
        delete c;
        *out = wrapped;
    });
}

extern "C" int bm25mi_corpus_view(const bm25mi_corpus *c, int64_t *num_docs, int64_t *vocab, int64_t *postings,
                                  const int64_t **term_offsets, const int32_t **doc_ids, const float **tf,
                                  const int32_t **doc_len)
{
    return Invoke([&] {
        Require(c != nullptr, "Null corpus");
        if (num_docs)
            *num_docs = c->docs;
        if (vocab)
            *vocab = c->vocab;
        if (postings)
            *postings = c->postings;
        if (term_offsets)
            *term_offsets = c->offsets.data();
        if (doc_ids)
            *doc_ids = c->ids.data();
        if (tf)
            *tf = c->tf.data();
        if (doc_len)
            *doc_len = c->len.data();
    });
}

extern "C" void bm25mi_corpus_free(bm25mi_corpus *c) { delete c; }

extern "C" int bm25mi_bm25_weights(const int64_t *term_offsets, int64_t vocab, const int32_t *doc_len, int64_t docs,
                                   double k1, double b, float *out_idf, float *out_norm, int32_t threads)
{
    return Invoke([&] {
        Require(term_offsets && doc_len && out_idf && out_norm, "Null weight arrays");
        Require(vocab >= 1 && docs >= 1, "Empty vocabulary or corpus");
        Require(std::isfinite(k1) && k1 > 0 && std::isfinite(b) && b >= 0 && b <= 1, "k1 > 0 and b in [0,1] required");
        double total = 0;
        for (int64_t d = 0; d < docs; ++d) {
            Require(doc_len[d] >= 0, "Negative document length");
            total += static_cast<double>(doc_len[d]);
        }
        Require(total > 0, "All-empty corpus unsupported");
        const double n = static_cast<double>(docs);
        const double average = total / n;
        const int th = Threads(threads);
        ParallelChunks(vocab, th, [&](int, int64_t lo, int64_t hi) {
            for (int64_t t = lo; t < hi; ++t) {
                const double df = static_cast<double>(term_offsets[t + 1] - term_offsets[t]);
                out_idf[t] = static_cast<float>(std::log1p((n - df + 0.5) / (df + 0.5)));
            }
        });
        ParallelChunks(docs, th, [&](int, int64_t lo, int64_t hi) {
            for (int64_t d = lo; d < hi; ++d)
                out_norm[d] = static_cast<float>(k1 * (1 - b + b * static_cast<double>(doc_len[d]) / average));
        });
    });
}
