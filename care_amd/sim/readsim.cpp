// readsim: seeded synthetic Illumina-like read generator and base-level
// evaluator for the CARE correction workloads.
//
//   readsim gen --profile {uniform-se|repeat-pe|mixed-se|large-pe}
//               --seed S [--genome N] [--coverage C] [--length L]
//               --out reads.fastq --truth truth.fastq
//   readsim eval --orig reads.fastq --corr corrected.fastq --truth truth.fastq
//
// Profiles (genome / reads / errors / qualities):
//   uniform-se : i.i.d. uniform ACGT genome (default 1 Mbp); single-end reads
//                of fixed length (default 100) at coverage 30, uniform start,
//                random strand; 1% uniform substitutions; phred 35 on correct
//                bases and 12 on substituted ones.
//   repeat-pe  : uniform genome (default 5 Mbp) with 10 repeat families (copy
//                length U[1,5] kbp, 5-20 copies, 1% divergence per copy);
//                2x150 paired reads at coverage 30, insert ~ N(300,30), mates
//                interleaved (pair k -> reads 2k, 2k+1); substitution rate
//                rising linearly 0.1% -> 2% along the read; phred =
//                round(-10 log10 p) + N(0,2) clipped to [2,41].
//   mixed-se   : 40% GC genome (default 20 Mbp) with 3% repeat content;
//                single-end reads of length U[100,250] at coverage 50;
//                0.5% of reads carry one N; per-read substitution rate
//                U[0.5%,2%]; phred 35 / 12.
//   large-pe   : uniform genome (default 100 Mbp) with 5% repeat content at
//                2% divergence; 2x150 paired reads at coverage 40, insert ~
//                N(350,40); substitution rate rising 0.1% -> 2.5% along the
//                read; phred 35 / 12.
//
// The truth file holds, for every read, the error-free sequence in the same
// orientation as the read (N positions keep the genome base).
//
// The evaluator counts, per base, TP (orig!=truth, corr==truth),
// FP (orig==truth, corr!=truth), FN (orig!=truth, corr!=truth) and
// TN (otherwise), the definitions of evaluationtool/arteval.cpp:183-203,
// and gain = (TP-FP)/(TP+FN) (arteval.cpp:227).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct Rng {
    std::uint64_t s[4];
    explicit Rng(std::uint64_t seed) {
        std::uint64_t z = seed;
        for (auto& x : s) {
            z += 0x9e3779b97f4a7c15ull;
            std::uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ull;
            t = (t ^ (t >> 27)) * 0x94d049bb133111ebull;
            x = t ^ (t >> 31);
        }
    }
    static std::uint64_t rotl(std::uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    std::uint64_t next() {
        const std::uint64_t result = rotl(s[1] * 5, 7) * 9;
        const std::uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return result;
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    std::uint64_t below(std::uint64_t n) { return n == 0 ? 0 : next() % n; }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

char complement(char c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        case 'T': return 'A';
        default: return c;
    }
}

std::string revcomp(const std::string& s) {
    std::string r(s.size(), 'N');
    for (std::size_t i = 0; i < s.size(); i++) r[s.size() - 1 - i] = complement(s[i]);
    return r;
}

char randomBase(Rng& rng, double gc) {
    const double u = rng.uniform();
    if (u < gc / 2) return 'C';
    if (u < gc) return 'G';
    if (u < gc + (1 - gc) / 2) return 'A';
    return 'T';
}

char substitute(Rng& rng, char c) {
    for (;;) {
        const char b = kBases[rng.below(4)];
        if (b != c) return b;
    }
}

// Inserts repeat families: each family is a random template; each copy is the
// template with per-base divergence, written over a random genome position.
void addRepeats(Rng& rng, std::string& genome, int families, int minLen, int maxLen,
                int minCopies, int maxCopies, double divergence, double gc) {
    for (int f = 0; f < families; f++) {
        const int len = minLen + int(rng.below(maxLen - minLen + 1));
        if (std::size_t(len) >= genome.size()) continue;
        std::string tmpl(len, 'A');
        for (auto& c : tmpl) c = randomBase(rng, gc);
        const int copies = minCopies + int(rng.below(maxCopies - minCopies + 1));
        for (int c = 0; c < copies; c++) {
            const std::size_t pos = rng.below(genome.size() - len);
            for (int i = 0; i < len; i++) {
                char b = tmpl[i];
                if (rng.uniform() < divergence) b = substitute(rng, b);
                genome[pos + i] = b;
            }
        }
    }
}

// Repeat content given as a fraction of the genome: families of copy length
// U[1,5] kbp with 5-20 copies are added until the fraction is reached.
void addRepeatFraction(Rng& rng, std::string& genome, double fraction, double divergence, double gc) {
    const double target = fraction * double(genome.size());
    if (genome.size() <= 1000) return;  // no copy of 1..5 kbp fits
    double covered = 0;
    while (covered < target) {
        const int len = 1000 + int(rng.below(4001));
        if (std::size_t(len) >= genome.size()) continue;  // copies shorter than the genome
        const int copies = 5 + int(rng.below(16));
        std::string tmpl(len, 'A');
        for (auto& c : tmpl) c = randomBase(rng, gc);
        for (int c = 0; c < copies && covered < target; c++) {
            const std::size_t pos = rng.below(genome.size() - len);
            for (int i = 0; i < len; i++) {
                char b = tmpl[i];
                if (rng.uniform() < divergence) b = substitute(rng, b);
                genome[pos + i] = b;
            }
            covered += len;
        }
    }
}

struct Writer {
    std::ofstream reads, truth;
    std::string buf, tbuf;
    Writer(const std::string& r, const std::string& t) : reads(r), truth(t) {
        if (!reads || !truth) throw std::runtime_error("cannot open output files");
    }
    void put(const std::string& name, const std::string& seq, const std::string& qual, const std::string& tru) {
        buf += '@'; buf += name; buf += '\n'; buf += seq; buf += "\n+\n"; buf += qual; buf += '\n';
        tbuf += '@'; tbuf += name; tbuf += '\n'; tbuf += tru; tbuf += "\n+\n"; tbuf += qual; tbuf += '\n';
        if (buf.size() > (1u << 22)) flush();
    }
    void flush() {
        reads.write(buf.data(), buf.size());
        truth.write(tbuf.data(), tbuf.size());
        buf.clear(); tbuf.clear();
    }
    ~Writer() { flush(); }
};

// Applies substitutions with a per-position probability; returns read and qualities.
template <class ProbFn, class QualFn>
void mutate(Rng& rng, const std::string& tru, std::string& read, std::string& qual, ProbFn prob, QualFn qualOf) {
    read = tru;
    qual.assign(tru.size(), 'I');
    for (std::size_t i = 0; i < tru.size(); i++) {
        const double p = prob(int(i));
        bool err = rng.uniform() < p;
        if (err) read[i] = substitute(rng, read[i]);
        qual[i] = qualOf(int(i), err, p);
    }
}

int run_gen(int argc, char** argv) {
    std::string profile, out, truthOut;
    std::uint64_t seed = 1;
    double genomeSize = -1, coverage = -1;
    int length = -1;
    double nfrac = 0.0;  // paired-end profiles: fraction of reads given 1-3 N bases
    for (int i = 2; i < argc; i++) {
        std::string a = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argc) throw std::runtime_error("missing value for " + a);
            return argv[++i];
        };
        if (a == "--profile") profile = val();
        else if (a == "--seed") seed = std::stoull(val());
        else if (a == "--genome") genomeSize = std::stod(val());
        else if (a == "--coverage") coverage = std::stod(val());
        else if (a == "--length") length = std::stoi(val());
        else if (a == "--nfrac") nfrac = std::stod(val());
        else if (a == "--out") out = val();
        else if (a == "--truth") truthOut = val();
        else throw std::runtime_error("unknown argument " + a);
    }
    if (out.empty() || truthOut.empty()) throw std::runtime_error("--out and --truth are required");

    Rng rng(seed);
    Writer w(out, truthOut);

    auto makeGenome = [&](std::size_t n, double gc) {
        std::string g(n, 'A');
        for (auto& c : g) c = randomBase(rng, gc);
        return g;
    };

    if (profile == "uniform-se") {
        const std::size_t G = genomeSize > 0 ? std::size_t(genomeSize) : 1000000;
        const double cov = coverage > 0 ? coverage : 30.0;
        const int L = length > 0 ? length : 100;
        std::string genome = makeGenome(G, 0.5);
        const std::size_t n = std::size_t(std::llround(cov * double(G) / L));
        std::string read, qual;
        for (std::size_t r = 0; r < n; r++) {
            const std::size_t pos = rng.below(G - L + 1);
            std::string tru = genome.substr(pos, L);
            if (rng.below(2)) tru = revcomp(tru);
            mutate(rng, tru, read, qual, [](int) { return 0.01; },
                   [](int, bool err, double) { return char(33 + (err ? 12 : 35)); });
            w.put("read" + std::to_string(r), read, qual, tru);
        }
    } else if (profile == "repeat-pe" || profile == "large-pe") {
        const bool large = profile == "large-pe";
        const std::size_t G = genomeSize > 0 ? std::size_t(genomeSize) : (large ? 100000000 : 5000000);
        const double cov = coverage > 0 ? coverage : (large ? 40.0 : 30.0);
        const int L = length > 0 ? length : 150;
        const double insMean = large ? 350 : 300, insSd = large ? 40 : 30;
        const double pLo = 0.001, pHi = large ? 0.025 : 0.02;
        std::string genome = makeGenome(G, 0.5);
        if (large) addRepeatFraction(rng, genome, 0.05, 0.02, 0.5);
        else addRepeats(rng, genome, 10, 1000, 5000, 5, 20, 0.01, 0.5);
        const std::size_t pairs = std::size_t(std::llround(cov * double(G) / (2.0 * L)));
        auto prob = [&](int i) { return pLo + (pHi - pLo) * double(i) / double(L - 1); };
        std::string read, qual;
        for (std::size_t p = 0; p < pairs; p++) {
            int ins = int(std::lround(insMean + insSd * rng.normal()));
            if (ins < L) ins = L;
            if (std::size_t(ins) > G) ins = int(G);
            const std::size_t pos = rng.below(G - ins + 1);
            std::string frag = genome.substr(pos, ins);
            if (rng.below(2)) frag = revcomp(frag);
            const std::string t1 = frag.substr(0, L);
            const std::string t2 = revcomp(frag.substr(ins - L, L));
            for (int m = 0; m < 2; m++) {
                const std::string& tru = m == 0 ? t1 : t2;
                if (large) {
                    mutate(rng, tru, read, qual, prob,
                           [](int, bool err, double) { return char(33 + (err ? 12 : 35)); });
                } else {
                    mutate(rng, tru, read, qual, prob, [&](int, bool, double pp) {
                        long q = std::lround(-10.0 * std::log10(pp) + 2.0 * rng.normal());
                        if (q < 2) q = 2;
                        if (q > 41) q = 41;
                        return char(33 + q);
                    });
                }
                if (nfrac > 0 && rng.uniform() < nfrac) {
                    const int count = 1 + int(rng.below(3));
                    for (int j = 0; j < count; j++) {
                        const int npos = int(rng.below(L));
                        read[npos] = 'N';
                        qual[npos] = char(33 + 2);
                    }
                }
                w.put("pair" + std::to_string(p) + "/" + std::to_string(m + 1), read, qual, tru);
            }
        }
    } else if (profile == "mixed-se") {
        const std::size_t G = genomeSize > 0 ? std::size_t(genomeSize) : 20000000;
        const double cov = coverage > 0 ? coverage : 50.0;
        const int Lmin = 100, Lmax = length > 0 ? length : 250;
        std::string genome = makeGenome(G, 0.4);
        addRepeatFraction(rng, genome, 0.03, 0.01, 0.4);
        const double meanLen = 0.5 * (Lmin + Lmax);
        const std::size_t n = std::size_t(std::llround(cov * double(G) / meanLen));
        std::string read, qual;
        for (std::size_t r = 0; r < n; r++) {
            const int L = Lmin + int(rng.below(Lmax - Lmin + 1));
            const std::size_t pos = rng.below(G - L + 1);
            std::string tru = genome.substr(pos, L);
            if (rng.below(2)) tru = revcomp(tru);
            const double rate = 0.005 + 0.015 * rng.uniform();
            mutate(rng, tru, read, qual, [&](int) { return rate; },
                   [](int, bool err, double) { return char(33 + (err ? 12 : 35)); });
            if (rng.uniform() < 0.005) {
                const int npos = int(rng.below(L));
                read[npos] = 'N';
                qual[npos] = char(33 + 2);
            }
            w.put("read" + std::to_string(r), read, qual, tru);
        }
    } else {
        throw std::runtime_error("unknown profile '" + profile + "'");
    }
    return 0;
}

struct FastqIn {
    std::ifstream in;
    explicit FastqIn(const std::string& f) : in(f) {
        if (!in) throw std::runtime_error("cannot open " + f);
    }
    bool next(std::string& seq) {
        std::string h, plus, q;
        if (!std::getline(in, h)) return false;
        std::getline(in, seq);
        std::getline(in, plus);
        std::getline(in, q);
        return true;
    }
};

int run_eval(int argc, char** argv) {
    std::string o, c, t;
    for (int i = 2; i + 1 < argc; i += 2) {
        std::string a = argv[i];
        if (a == "--orig") o = argv[i + 1];
        else if (a == "--corr") c = argv[i + 1];
        else if (a == "--truth") t = argv[i + 1];
        else throw std::runtime_error("unknown argument " + a);
    }
    FastqIn fo(o), fc(c), ft(t);
    std::string so, sc, st;
    std::uint64_t TP = 0, FP = 0, FN = 0, TN = 0, reads = 0;
    while (fo.next(so) && fc.next(sc) && ft.next(st)) {
        reads++;
        const std::size_t n = std::min(so.size(), std::min(sc.size(), st.size()));
        for (std::size_t i = 0; i < n; i++) {
            const bool origOk = so[i] == st[i];
            const bool corrOk = sc[i] == st[i];
            if (!origOk && corrOk) TP++;
            else if (origOk && !corrOk) FP++;
            else if (!origOk && !corrOk) FN++;
            else TN++;
        }
    }
    const double gain = (TP + FN) ? (double(TP) - double(FP)) / double(TP + FN) : 0.0;
    std::printf("{\"reads\": %llu, \"TP\": %llu, \"FP\": %llu, \"FN\": %llu, \"TN\": %llu, \"gain\": %.6f}\n",
                (unsigned long long)reads, (unsigned long long)TP, (unsigned long long)FP,
                (unsigned long long)FN, (unsigned long long)TN, gain);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc < 2) throw std::runtime_error("usage: readsim gen|eval ...");
        const std::string mode = argv[1];
        if (mode == "gen") return run_gen(argc, argv);
        if (mode == "eval") return run_eval(argc, argv);
        throw std::runtime_error("unknown mode " + mode);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "readsim: %s\n", e.what());
        return 1;
    }
}
