// vcf.cpp -- VCF reader of the kestrel CLI (vcf.rs:96-482 behaviour).
#include "vcf.hpp"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace kvcf {

namespace {

struct Reader {
    gzFile f = nullptr;
    std::vector<char> buf = std::vector<char>(1 << 20);
    ~Reader() {
        if (f) gzclose(f);
    }
    bool line(std::string& out) {
        out.clear();
        for (;;) {
            if (!gzgets(f, buf.data(), (int)buf.size())) return !out.empty();
            size_t n = strlen(buf.data());
            out.append(buf.data(), n);
            if (n && out.back() == '\n') {
                out.pop_back();
                if (!out.empty() && out.back() == '\r') out.pop_back();
                return true;
            }
        }
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
    std::string s(b, e);
    char* end = nullptr;
    float v = strtof(s.c_str(), &end);
    if (end == s.c_str() || *end) *ok = false;
    return (double)v;
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

bool read(const std::string& path, Field field, bool parse_af, Data& out, std::string& err) {
    Reader r;
    r.f = gzopen(path.c_str(), "rb");
    if (!r.f) {
        err = "No such file or directory (os error 2): " + path;
        return false;
    }
    gzbuffer(r.f, 1 << 20);
    std::string line;
    bool header = false;
    std::vector<std::pair<const char*, const char*>> cols, fmt, sub, vals;
    const char* key = field == Field::GT ? "GT" : field == Field::GL ? "GL" : "PL";
    std::vector<uint8_t> codes;
    std::vector<double> site_probs;
    uint64_t recno = 0;
    while (r.line(line)) {
        if (line.empty()) continue;
        if (line.compare(0, 2, "##") == 0) continue;
        if (line[0] == '#') {
            split(line.data(), line.data() + line.size(), '\t', cols);
            out.samples.clear();
            for (size_t i = 9; i < cols.size(); i++) out.samples.push_back(str(cols[i]));
            header = true;
            continue;
        }
        if (!header) {
            err = "invalid header: missing #CHROM line";
            return false;
        }
        recno++;
        split(line.data(), line.data() + line.size(), '\t', cols);
        const size_t n = out.samples.size();
        if (cols.size() < 8 || (n && cols.size() != 9 + n)) {
            err = "Error reading record " + std::to_string(recno) + ": wrong number of columns";
            return false;
        }
        out.total++;
        const int nb = count_snp_bases(cols[3], cols[4]);
        if (nb < 2) {
            out.not_snp++;
            continue;
        }
        if (nb > 2) {
            out.multiallelic++;
            err = "record " + std::to_string(recno) + " (" + str(cols[0]) + ":" + str(cols[1]) +
                  ") has " + std::to_string(nb) + " alleles: tri- and tetra-allelic sites are not supported by this "
                  "build";
            return false;
        }
        int kidx = -1;
        if (n) {
            split(cols[8].first, cols[8].second, ':', fmt);
            for (size_t k = 0; k < fmt.size(); k++)
                if ((size_t)(fmt[k].second - fmt[k].first) == strlen(key) && !strncmp(fmt[k].first, key, strlen(key)))
                    kidx = (int)k;
            if (kidx < 0) {
                err = std::string("No ") + key + " data found";
                return false;
            }
        }
        codes.assign(n, 3);
        if (field != Field::GT) site_probs.assign(3 * n, NAN);
        bool skip = false;
        for (size_t i = 0; i < n && !skip; i++) {
            split(cols[9 + i].first, cols[9 + i].second, ':', sub);
            if ((size_t)kidx >= sub.size()) {
                err = field == Field::GT ? "No genotype for sample found" : "No genotype likelihood for sample found";
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
                        char* end = nullptr;
                        std::string tok(s, p);
                        long a = strtol(tok.c_str(), &end, 10);
                        if (tok.empty() || *end || a < 0) {
                            err = "Error reading genotype";
                            return false;
                        }
                        if (a > 255) {
                            err = "Allele is greater than 255";
                            return false;
                        }
                        if (ng >= 2) {
                            err = "Genotype is not diploid!";
                            return false;
                        }
                        al[ng++] = (int)a;
                        s = p + 1;
                    }
                }
                if (skip) break;
                if (ng != 2) {
                    err = "Genotype is not diploid!";
                    return false;
                }
                const int lo = al[0] < al[1] ? al[0] : al[1], hi = al[0] < al[1] ? al[1] : al[0];
                if (hi >= nb) {
                    err = "allele index " + std::to_string(hi) + " exceeds the alleles of record " +
                          std::to_string(recno);
                    return false;
                }
                codes[i] = (uint8_t)(hi * (hi + 1) / 2 + lo);  // coefficients.rs:27-32
            } else {
                if (e - b == 1 && *b == '.') {
                    err = "No genotype likelihood for sample found";
                    return false;
                }
                split(b, e, ',', vals);
                if (vals.size() != 3) {
                    err = vals.size() > 3 ? "Too many genotype likelihoods!" : "Not enough genotype likelihoods!";
                    return false;
                }
                double v[3], prob[3];
                bool any_nan = false;
                for (int k = 0; k < 3; k++) {
                    bool miss, ok;
                    if (field == Field::GL) {
                        v[k] = parse_f32(vals[k].first, vals[k].second, &miss, &ok);
                    } else {
                        miss = (vals[k].second - vals[k].first == 1 && *vals[k].first == '.');
                        ok = true;
                        if (!miss) {
                            std::string tok(vals[k].first, vals[k].second);
                            char* end = nullptr;
                            long pl = strtol(tok.c_str(), &end, 10);
                            ok = !tok.empty() && !*end;
                            v[k] = (double)pl;
                        } else
                            v[k] = NAN;
                    }
                    if (!ok) {
                        err = "Error reading genotype likelihood";
                        return false;
                    }
                }
                if (field == Field::GL) {  // vcf.rs:349-368: normalise by the maximum GL
                    double mx = -INFINITY;
                    for (int k = 0; k < 3; k++) mx = std::fmax(mx, v[k]);
                    for (int k = 0; k < 3; k++) prob[k] = std::pow(10.0, v[k] - mx);
                } else {  // vcf.rs:407
                    for (int k = 0; k < 3; k++) prob[k] = std::pow(10.0, -v[k] / 10.0);
                }
                int one = -1, zeros = 0;
                for (int k = 0; k < 3; k++) {
                    if (prob[k] != prob[k]) any_nan = true;
                    else if (prob[k] == 1.0 && one < 0) one = k;
                    else if (prob[k] == 0.0) zeros++;
                }
                for (int k = 0; k < 3; k++) site_probs[3 * i + k] = prob[k];
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
            continue;
        }
        if (parse_af) {
            // vcf.rs:446-482: AF INFO -> [1 - clamp(sum alt, 0, 1), alt...]
            split(cols[7].first, cols[7].second, ';', sub);
            bool found = false;
            for (auto& t : sub) {
                if (t.second - t.first > 3 && !strncmp(t.first, "AF=", 3)) {
                    split(t.first + 3, t.second, ',', vals);
                    if (vals.size() != 1) {
                        err = vals.size() > 1 ? "Too many allele frequencies!" : "Not enough allele frequencies!";
                        return false;
                    }
                    bool miss, ok;
                    const double a = parse_f32(vals[0].first, vals[0].second, &miss, &ok);
                    if (miss || !ok) {
                        err = miss ? "No AF data found" : "Error reading AF";
                        return false;
                    }
                    double alt_sum = 0.0;
                    alt_sum += a;
                    out.af.push_back(1.0 - std::fmin(std::fmax(alt_sum, 0.0), 1.0));
                    out.af.push_back(a);
                    found = true;
                    break;
                }
            }
            if (!found) {
                err = "No AF info found";
                return false;
            }
        }
        out.site_codes.insert(out.site_codes.end(), codes.begin(), codes.end());
        if (field != Field::GT) out.probs.insert(out.probs.end(), site_probs.begin(), site_probs.end());
        out.n_sites++;
    }
    if (!header) {
        err = "invalid header: missing #CHROM line";
        return false;
    }
    return true;
}

}  // namespace kvcf
