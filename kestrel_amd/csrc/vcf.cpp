// vcf.cpp -- VCF reader of the kestrel CLI (vcf.rs:96-482 behaviour).
//
// A pipeline over a thread pool: BGZF blocks are inflated in parallel (a plain
// gzip stream or plain text is read in order), the text is cut at line
// boundaries into batches of records, the batches are parsed in parallel and
// merged in file order, so the decoded cohort and the first error are those of
// a serial read for any number of threads.
#include "vcf.hpp"

#include <zlib.h>

#include <charconv>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <future>
#include <memory>
#include <mutex>
#include <thread>

namespace kvcf {

namespace {

class Pool {
    std::vector<std::thread> th_;
    std::deque<std::function<void()>> q_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stop_ = false;

public:
    explicit Pool(unsigned n) {
        for (unsigned i = 0; i < n; i++)
            th_.emplace_back([this] {
                for (;;) {
                    std::function<void()> f;
                    {
                        std::unique_lock<std::mutex> lk(mu_);
                        cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
                        if (q_.empty()) return;
                        f = std::move(q_.front());
                        q_.pop_front();
                    }
                    f();
                }
            });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    template <class F>
    auto submit(F f) -> std::future<decltype(f())> {
        auto task = std::make_shared<std::packaged_task<decltype(f())()>>(std::move(f));
        auto fut = task->get_future();
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.emplace_back([task] { (*task)(); });
        }
        cv_.notify_one();
        return fut;
    }
};

void split(const char* b, const char* e, char sep, std::vector<std::pair<const char*, const char*>>& out) {
    out.clear();
    const char* s = b;
    for (const char* p = b; p <= e; p++)
        if (p == e || *p == sep) {
            out.emplace_back(s, p);
            s = p + 1;
        }
}

bool is_base(const char* b, const char* e) {
    if (e - b != 1) return false;
    switch (*b) {
    case 'A': case 'C': case 'G': case 'T': case 'a': case 'c': case 'g': case 't': return true;
    default: return false;
    }
}

// vcf.rs:424-444: REF and every ALT a single A/C/G/T; returns the allele count
// (0 when the record is not a SNP)
int count_snp_bases(const std::pair<const char*, const char*>& ref, const std::pair<const char*, const char*>& alt) {
    if (!is_base(ref.first, ref.second)) return 0;
    if (alt.second - alt.first == 1 && *alt.first == '.') return 1;
    int n = 1;
    const char* s = alt.first;
    for (const char* p = alt.first; p <= alt.second; p++)
        if (p == alt.second || *p == ',') {
            if (!is_base(s, p)) return 0;
            n++;
            s = p + 1;
        }
    return n;
}

std::string str(const std::pair<const char*, const char*>& t) { return std::string(t.first, t.second); }

// GL / PL values are stored as 32-bit floats by the VCF library the reference
// uses, then widened (vcf.rs:355).
double parse_f32(const char* b, const char* e, bool* missing, bool* ok) {
    *missing = (e - b == 1 && *b == '.');
    *ok = true;
    if (*missing) return NAN;
    const char* p = b;
    if (p < e && *p == '+') p++;  // a leading '+' is accepted, as by Rust's f32 parser
    float v = 0.0f;
    const auto r = std::from_chars(p, e, v, std::chars_format::general);  // correctly rounded
    if (p == e || r.ptr != e) {
        *ok = false;
        return 0.0;
    }
    if (r.ec == std::errc::result_out_of_range) {  // overflow to inf, underflow to 0 (rare): strtof's result
        char buf[64];
        const size_t n = (size_t)(e - b);
        if (n >= sizeof buf) {
            *ok = false;
            return 0.0;
        }
        memcpy(buf, b, n);
        buf[n] = 0;
        return (double)strtof(buf, nullptr);
    }
    if (r.ec != std::errc()) *ok = false;
    return (double)v;
}

// pow(10, x) with libm's results where they are fixed: exactly 1 at x = 0 (C
// pow(x, +-0) = 1) and exactly 0 far below the smallest subnormal (10^-400);
// one-hot likelihood rows (0 and -1000, PL 0 and 10000) take only these.
inline double pow10(double x) {
    if (x == 0.0) return 1.0;
    if (x <= -400.0) return 0.0;
    return std::pow(10.0, x);
}

// Decimal digits only (a GT allele), with an optional leading '+' as strtol
// accepts; values are capped so that a long digit run cannot overflow.
bool parse_uint(const char* b, const char* e, long* v) {
    if (b < e && *b == '+') b++;
    if (b == e) return false;
    long x = 0;
    for (const char* p = b; p < e; p++) {
        if (*p < '0' || *p > '9') return false;
        if (x < 1000000) x = x * 10 + (*p - '0');
    }
    *v = x;
    return true;
}

// An optionally signed decimal integer (a PL value), as strtol over the whole token.
bool parse_int(const char* b, const char* e, long* v) {
    bool neg = false;
    if (b < e && (*b == '+' || *b == '-')) {
        neg = *b == '-';
        b++;
    }
    if (b == e) return false;
    long x = 0;
    for (const char* p = b; p < e; p++) {
        if (*p < '0' || *p > '9') return false;
        if (x < (1L << 40)) x = x * 10 + (*p - '0');
    }
    *v = neg ? -x : x;
    return true;
}


// ------------------------------------------------------------ text sources

struct Chunk {
    std::string text;
    std::string err;
};

// BGZF (the blocked gzip of htslib): every block is a gzip member with a 'BC'
// extra field holding its size; groups of blocks are inflated in parallel.
class BgzfSource {
    FILE* f_;
    Pool& pool_;
    std::deque<std::future<Chunk>> pending_;
    bool eof_ = false;
    size_t ahead_;
    // a block: 18 header bytes (with the 'BC' field), deflate data, CRC32 and
    // ISIZE (8 bytes); BSIZE + 1 is its length, ISIZE at most 65,536
    static constexpr size_t BLOCK_MIN = 18 + 8;
    static constexpr size_t BLOCK_MAX = 65536;

    static Chunk inflate_group(std::vector<std::string> blocks) {
        Chunk c;
        for (const std::string& b : blocks) {
            const size_t isize = (size_t)(uint8_t)b[b.size() - 4] | (size_t)(uint8_t)b[b.size() - 3] << 8 |
                                 (size_t)(uint8_t)b[b.size() - 2] << 16 | (size_t)(uint8_t)b[b.size() - 1] << 24;
            if (isize > BLOCK_MAX) {  // a block inflates to at most 64 KiB
                c.err = "bgzf: corrupt block";
                return c;
            }
            const size_t off = c.text.size();
            c.text.resize(off + isize);
            z_stream s;
            memset(&s, 0, sizeof s);
            if (inflateInit2(&s, -15) != Z_OK) {
                c.err = "bgzf: inflateInit failed";
                return c;
            }
            s.next_in = (Bytef*)(b.data() + 18);
            s.avail_in = (uInt)(b.size() - 18 - 8);
            s.next_out = (Bytef*)(&c.text[off]);
            s.avail_out = (uInt)isize;
            const int rc = inflate(&s, Z_FINISH);
            inflateEnd(&s);
            if (rc != Z_STREAM_END || s.avail_out != 0) {
                c.err = "bgzf: corrupt block";
                return c;
            }
        }
        return c;
    }

public:
    BgzfSource(FILE* f, Pool& pool, size_t ahead) : f_(f), pool_(pool), ahead_(ahead) {}
    ~BgzfSource() {
        for (auto& p : pending_) p.wait();
        fclose(f_);
    }
    static bool is_bgzf(const unsigned char* h, size_t n) {
        return n >= 18 && h[0] == 0x1f && h[1] == 0x8b && h[2] == 8 && (h[3] & 4) && h[12] == 'B' && h[13] == 'C';
    }
    bool next(Chunk& out) {
        while (!eof_ && pending_.size() < ahead_) {
            std::vector<std::string> group;
            size_t bytes = 0;
            while (bytes < (1u << 20)) {  // ~1 MB of compressed blocks per job
                unsigned char h[18];
                const size_t got = fread(h, 1, 18, f_);
                if (got == 0) {
                    eof_ = true;
                    break;
                }
                if (got < 18 || !is_bgzf(h, got)) {
                    out.err = "bgzf: bad block header";
                    return true;
                }
                const size_t bsize = (size_t)h[16] | (size_t)h[17] << 8;
                if (bsize + 1 < BLOCK_MIN) {  // shorter than its own header and footer
                    out.err = "bgzf: bad block size";
                    return true;
                }
                std::string blk(bsize + 1, '\0');
                memcpy(&blk[0], h, 18);
                if (fread(&blk[18], 1, bsize + 1 - 18, f_) != bsize + 1 - 18) {
                    out.err = "bgzf: truncated block";
                    return true;
                }
                bytes += blk.size();
                group.push_back(std::move(blk));
            }
            if (!group.empty()) pending_.push_back(pool_.submit([g = std::move(group)]() mutable {
                return inflate_group(std::move(g));
            }));
        }
        if (pending_.empty()) return false;
        out = pending_.front().get();
        pending_.pop_front();
        return true;
    }
};

// Plain gzip (a single stream) or uncompressed text, in order.
class GzSource {
    gzFile f_;

public:
    explicit GzSource(gzFile f) : f_(f) { gzbuffer(f_, 1 << 20); }
    ~GzSource() { gzclose(f_); }
    bool next(Chunk& out) {
        out.text.resize(4u << 20);
        const int n = gzread(f_, &out.text[0], (unsigned)out.text.size());
        if (n < 0) {
            int e = 0;
            out.err = std::string("gzip: ") + gzerror(f_, &e);
            out.text.clear();
            return true;
        }
        out.text.resize((size_t)n);
        return n > 0;
    }
};

// ------------------------------------------------------------ records

struct Ctx {
    Field field;
    bool parse_af;
    const std::vector<std::string>* samples;
};

// The decoded records of one batch, in order.
struct Batch {
    std::vector<uint8_t> codes;  // kept sites x n
    std::vector<double> probs;   // kept sites x n x 3 (GL / PL)
    std::vector<double> af;      // kept sites x 2 (INFO AF)
    std::vector<double> probs3, probs4;  // GL / PL sites with 3 and 4 alleles
    std::vector<uint8_t> codes3, codes4; // GT sites with 3 and 4 alleles: genotype indices
    std::vector<double> af3, af4;        // their INFO AF (x 3, x 4)
    uint32_t kept = 0, kept3 = 0, kept4 = 0;
    uint64_t total = 0, not_snp = 0, skipped_missing = 0, multiallelic = 0;
    bool one_hot = true;
    std::string err;  // the first error of the batch (parsing stops there)
};

struct Scratch {
    std::vector<std::pair<const char*, const char*>> cols, fmt, sub, vals;
    std::vector<uint8_t> codes;
    std::vector<double> probs;
};

// One record (vcf.rs:96-482); false with b.err set on an error.
bool parse_record(const char* lb, const char* le, uint64_t recno, const Ctx& ctx, Batch& out, Scratch& sc) {
    auto& cols = sc.cols;
    auto& fmt = sc.fmt;
    auto& sub = sc.sub;
    auto& vals = sc.vals;
    const Field field = ctx.field;
    split(lb, le, '\t', cols);
    const size_t n = ctx.samples->size();
    if (cols.size() < 8 || (n && cols.size() != 9 + n)) {
        out.err = "Error reading record " + std::to_string(recno) + ": wrong number of columns";
        return false;
    }
    out.total++;
    const int nb = count_snp_bases(cols[3], cols[4]);
    if (nb < 2) {
        out.not_snp++;
        return true;
    }
    const int ng = nb * (nb + 1) / 2;  // likelihoods per sample (vcf.rs:270-305)
    if (nb > 2) out.multiallelic++;
    if (nb > 2 && ctx.parse_af && field != Field::GT) {  // wrapper.rs:113-122 leaves GL with given frequencies unimplemented
        out.err = "record " + std::to_string(recno) + ": INFO allele frequencies with tri- and tetra-allelic "
                  "likelihoods are not supported";
        return false;
    }
    const char* key = field == Field::GT ? "GT" : field == Field::GL ? "GL" : "PL";
    int kidx = -1;
    if (n) {
        split(cols[8].first, cols[8].second, ':', fmt);
        for (size_t k = 0; k < fmt.size(); k++)
            if ((size_t)(fmt[k].second - fmt[k].first) == strlen(key) && !strncmp(fmt[k].first, key, strlen(key)))
                kidx = (int)k;
        if (kidx < 0) {
            out.err = std::string("No ") + key + " data found";
            return false;
        }
    }
    auto& codes = sc.codes;
    auto& site_probs = sc.probs;
    codes.assign(n, 3);
    if (field != Field::GT) site_probs.assign((size_t)ng * n, NAN);
    bool skip = false;
    for (size_t i = 0; i < n && !skip; i++) {
        split(cols[9 + i].first, cols[9 + i].second, ':', sub);
        if ((size_t)kidx >= sub.size()) {
            out.err = field == Field::GT ? "No genotype for sample found" : "No genotype likelihood for sample found";
            return false;
        }
        const char* b = sub[kidx].first;
        const char* e = sub[kidx].second;
        if (field == Field::GT) {
            // vcf.rs:195-236
            int al[2] = {0, 0}, ng = 0;
            const char* s = b;
            for (const char* p = b; p <= e && !skip; p++) {
                if (p == e || *p == '/' || *p == '|') {
                    if (p - s == 1 && *s == '.') {
                        skip = true;  // a missing allele drops the site
                        break;
                    }
                    long a = 0;
                    if (!parse_uint(s, p, &a)) {
                        out.err = "Error reading genotype";
                        return false;
                    }
                    if (a > 255) {
                        out.err = "Allele is greater than 255";
                        return false;
                    }
                    if (ng >= 2) {
                        out.err = "Genotype is not diploid!";
                        return false;
                    }
                    al[ng++] = (int)a;
                    s = p + 1;
                }
            }
            if (skip) break;
            if (ng != 2) {
                out.err = "Genotype is not diploid!";
                return false;
            }
            const int lo = al[0] < al[1] ? al[0] : al[1], hi = al[0] < al[1] ? al[1] : al[0];
            if (hi >= nb) {
                out.err = "allele index " + std::to_string(hi) + " exceeds the alleles of record " +
                          std::to_string(recno);
                return false;
            }
            codes[i] = (uint8_t)(hi * (hi + 1) / 2 + lo);  // coefficients.rs:27-32
        } else {
            if (e - b == 1 && *b == '.') {
                out.err = "No genotype likelihood for sample found";
                return false;
            }
            split(b, e, ',', vals);
            if (vals.size() != (size_t)ng) {
                out.err = vals.size() > (size_t)ng ? "Too many genotype likelihoods!" : "Not enough genotype likelihoods!";
                return false;
            }
            double v[10], prob[10];
            bool any_nan = false;
            for (int k = 0; k < ng; k++) {
                bool miss, ok;
                if (field == Field::GL) {
                    v[k] = parse_f32(vals[k].first, vals[k].second, &miss, &ok);
                } else {
                    miss = (vals[k].second - vals[k].first == 1 && *vals[k].first == '.');
                    ok = true;
                    if (!miss) {
                        long pl = 0;
                        ok = parse_int(vals[k].first, vals[k].second, &pl);
                        v[k] = (double)pl;
                    } else
                        v[k] = NAN;
                }
                if (!ok) {
                    out.err = "Error reading genotype likelihood";
                    return false;
                }
            }
            if (field == Field::GL) {  // vcf.rs:349-368: normalise by the maximum GL
                double mx = -INFINITY;
                for (int k = 0; k < ng; k++) mx = std::fmax(mx, v[k]);
                for (int k = 0; k < ng; k++) prob[k] = pow10(v[k] - mx);
            } else {  // vcf.rs:407
                for (int k = 0; k < ng; k++) prob[k] = pow10(-v[k] / 10.0);
            }
            for (int k = 0; k < ng; k++) site_probs[(size_t)ng * i + k] = prob[k];
            if (ng != 3) continue;  // 3 / 4 alleles: the likelihood rows are all the path reads
            int one = -1, zeros = 0;
            for (int k = 0; k < 3; k++) {
                if (prob[k] != prob[k]) any_nan = true;
                else if (prob[k] == 1.0 && one < 0) one = k;
                else if (prob[k] == 0.0) zeros++;
            }
            if (any_nan) {
                codes[i] = 3;  // the row is dropped for every pair with this sample
            } else if (one >= 0 && zeros == 2) {
                codes[i] = (uint8_t)one;
            } else {
                codes[i] = 0;  // a soft row: only the likelihood path reads it
                out.one_hot = false;
            }
        }
    }
    if (skip) {
        out.skipped_missing++;
        return true;
    }
    if (ctx.parse_af) {
        // vcf.rs:446-482: AF INFO -> [1 - clamp(sum alt, 0, 1), alt...]
        split(cols[7].first, cols[7].second, ';', sub);
        bool found = false;
        for (auto& t : sub) {
            if (t.second - t.first > 3 && !strncmp(t.first, "AF=", 3)) {
                split(t.first + 3, t.second, ',', vals);
                double afv[4] = {0.0, 0.0, 0.0, 0.0}, alt_sum = 0.0;
                int num_af = 1;
                for (auto& v : vals) {
                    if (num_af >= nb) {
                        out.err = "Too many allele frequencies!";
                        return false;
                    }
                    bool miss, ok;
                    const double a = parse_f32(v.first, v.second, &miss, &ok);
                    if (miss || !ok) {
                        out.err = miss ? "No AF data found" : "Error reading AF";
                        return false;
                    }
                    alt_sum += a;
                    afv[num_af++] = a;
                }
                if (num_af != nb) {
                    out.err = "Not enough allele frequencies!";
                    return false;
                }
                afv[0] = 1.0 - std::fmin(std::fmax(alt_sum, 0.0), 1.0);
                std::vector<double>& dst = nb == 2 ? out.af : nb == 3 ? out.af3 : out.af4;
                dst.insert(dst.end(), afv, afv + nb);
                found = true;
                break;
            }
        }
        if (!found) {
            out.err = "No AF info found";
            return false;
        }
    }
    if (nb == 3) {
        if (field == Field::GT) out.codes3.insert(out.codes3.end(), codes.begin(), codes.end());
        else out.probs3.insert(out.probs3.end(), site_probs.begin(), site_probs.end());
        out.kept3++;
        return true;
    }
    if (nb == 4) {
        if (field == Field::GT) out.codes4.insert(out.codes4.end(), codes.begin(), codes.end());
        else out.probs4.insert(out.probs4.end(), site_probs.begin(), site_probs.end());
        out.kept4++;
        return true;
    }
    out.codes.insert(out.codes.end(), codes.begin(), codes.end());
    if (field != Field::GT) out.probs.insert(out.probs.end(), site_probs.begin(), site_probs.end());
    out.kept++;
    return true;
}

// Parses the records of text (complete lines), numbering from first_recno.
Batch parse_batch(const std::string& text, uint64_t first_recno, const Ctx& ctx) {
    Batch b;
    Scratch sc;
    uint64_t recno = first_recno;
    const char* p = text.data();
    const char* end = p + text.size();
    while (p < end) {
        const char* nl = (const char*)memchr(p, '\n', (size_t)(end - p));
        const char* le = nl ? nl : end;
        const char* lt = le;
        if (lt > p && lt[-1] == '\r') lt--;
        if (lt > p && !(lt - p >= 2 && p[0] == '#' && p[1] == '#')) {
            if (p[0] == '#') {
                b.err = "Error reading record " + std::to_string(recno) + ": a header line among the records";
                return b;
            }
            if (!parse_record(p, lt, recno, ctx, b, sc)) return b;
            recno++;
        }
        p = le + 1;
    }
    return b;
}

}  // namespace

std::vector<uint8_t> Data::pack(size_t* row_bytes) const {
    const size_t n = samples.size();
    const size_t rb = (n_sites + 3) / 4;
    *row_bytes = rb;
    std::vector<uint8_t> out(n * rb, 0xff);  // padding = code 3
    for (size_t s = 0; s < n_sites; s++) {
        const uint8_t* sc = site_codes.data() + s * n;
        const size_t byte = s >> 2, sh = (s & 3) * 2;
        for (size_t i = 0; i < n; i++) {
            uint8_t& b = out[i * rb + byte];
            b = (uint8_t)((b & ~(3u << sh)) | ((sc[i] & 3u) << sh));
        }
    }
    return out;
}

void Data::join_probs() {
    size_t total = probs.size();
    for (const auto& p : probs_parts) total += p.size();
    probs.reserve(total);
    for (auto& p : probs_parts) {
        probs.insert(probs.end(), p.begin(), p.end());
        std::vector<double>().swap(p);
    }
    probs_parts.clear();
}

std::vector<uint8_t> Data::sample_major(int alleles) const {
    const size_t n = samples.size();
    const std::vector<uint8_t>& src = alleles == 3 ? site_codes3 : site_codes4;
    const size_t v = alleles == 3 ? n_sites3 : n_sites4;
    std::vector<uint8_t> out(n * v);
    constexpr size_t B = 64;  // transpose in 64 x 64 tiles
    for (size_t s0 = 0; s0 < v; s0 += B)
        for (size_t i0 = 0; i0 < n; i0 += B)
            for (size_t s = s0; s < std::min(v, s0 + B); s++)
                for (size_t i = i0; i < std::min(n, i0 + B); i++) out[i * v + s] = src[s * n + i];
    return out;
}

bool read(const std::string& path, Field field, bool parse_af, Data& out, std::string& err, unsigned threads) {
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) {
        err = "No such file or directory (os error 2): " + path;
        return false;
    }
    unsigned char head[18];
    const size_t got = fread(head, 1, sizeof head, f);
    Pool pool(threads);
    std::unique_ptr<BgzfSource> bgzf;
    std::unique_ptr<GzSource> gz;
    if (BgzfSource::is_bgzf(head, got)) {
        rewind(f);
        bgzf.reset(new BgzfSource(f, pool, 2 * (size_t)threads));
    } else {
        fclose(f);
        gzFile g = gzopen(path.c_str(), "rb");
        if (!g) {
            err = "No such file or directory (os error 2): " + path;
            return false;
        }
        gz.reset(new GzSource(g));
    }
    auto next = [&](Chunk& c) { return bgzf ? bgzf->next(c) : gz->next(c); };

    const Ctx ctx{field, parse_af, &out.samples};
    bool header = false;
    uint64_t recno = 1;  // number of the next record
    std::deque<std::future<Batch>> pending;
    const size_t max_pending = 3 * (size_t)threads;
    bool failed = false;
    auto merge = [&](Batch&& b) {
        if (failed) return;
        out.total += b.total;
        out.not_snp += b.not_snp;
        out.skipped_missing += b.skipped_missing;
        out.multiallelic += b.multiallelic;
        out.one_hot = out.one_hot && b.one_hot;
        out.site_codes.insert(out.site_codes.end(), b.codes.begin(), b.codes.end());
        if (!b.probs.empty()) out.probs_parts.push_back(std::move(b.probs));  // joined on demand
        out.af.insert(out.af.end(), b.af.begin(), b.af.end());
        out.probs3.insert(out.probs3.end(), b.probs3.begin(), b.probs3.end());
        out.probs4.insert(out.probs4.end(), b.probs4.begin(), b.probs4.end());
        out.site_codes3.insert(out.site_codes3.end(), b.codes3.begin(), b.codes3.end());
        out.site_codes4.insert(out.site_codes4.end(), b.codes4.begin(), b.codes4.end());
        out.af3.insert(out.af3.end(), b.af3.begin(), b.af3.end());
        out.af4.insert(out.af4.end(), b.af4.begin(), b.af4.end());
        out.n_sites += b.kept;
        out.n_sites3 += b.kept3;
        out.n_sites4 += b.kept4;
        if (!b.err.empty()) {
            failed = true;
            err = b.err;
        }
    };
    auto submit = [&](std::string&& text, uint64_t first) {
        while (pending.size() >= max_pending) {
            merge(pending.front().get());
            pending.pop_front();
        }
        pending.push_back(pool.submit([t = std::move(text), first, &ctx]() { return parse_batch(t, first, ctx); }));
    };

    std::string carry, batch, scan_err;
    uint64_t batch_first = 1;
    std::vector<std::pair<const char*, const char*>> cols;
    Chunk c;
    bool more = true;
    while (more && scan_err.empty()) {
        more = next(c);
        if (!c.err.empty()) {
            scan_err = c.err;
            break;
        }
        if (more)
            carry.append(c.text);
        else if (!carry.empty() && carry.back() != '\n')
            carry.push_back('\n');  // a last line without a newline
        const size_t last = carry.rfind('\n');
        if (last == std::string::npos) continue;
        const char* p = carry.data();
        const char* end = carry.data() + last + 1;
        while (p < end) {
            const char* nl = (const char*)memchr(p, '\n', (size_t)(end - p));
            const char* lt = nl;
            if (lt > p && lt[-1] == '\r') lt--;
            const bool empty = lt == p, meta = lt - p >= 2 && p[0] == '#' && p[1] == '#';
            if (!header) {
                if (!empty && !meta) {
                    if (p[0] != '#') {
                        scan_err = "invalid header: missing #CHROM line";
                        break;
                    }
                    split(p, lt, '\t', cols);
                    out.samples.clear();
                    for (size_t k = 9; k < cols.size(); k++) out.samples.push_back(str(cols[k]));
                    header = true;
                }
            } else {
                if (batch.empty()) batch_first = recno;
                batch.append(p, (size_t)(nl + 1 - p));
                if (!empty && !meta) recno++;
                if (batch.size() >= (2u << 20)) {
                    submit(std::move(batch), batch_first);
                    batch.clear();
                }
            }
            p = nl + 1;
        }
        carry.erase(0, last + 1);
    }
    if (scan_err.empty() && !batch.empty()) submit(std::move(batch), batch_first);
    // records before a reading error come first, as in a serial read
    for (auto& fut : pending) merge(fut.get());
    pending.clear();
    if (failed) return false;
    if (!scan_err.empty()) {
        err = scan_err;
        return false;
    }
    if (!header) {
        err = "invalid header: missing #CHROM line";
        return false;
    }
    return true;
}

}  // namespace kvcf
