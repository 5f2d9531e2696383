// cli.cpp -- the `kestrel` command line (cli.rs:12-181): VCF in, per-pair
// kinship, Jacquard coefficients, iterations, objective and condition out as
// TSV, computed by libkestrel_amd on the GPU.
//
//   kestrel INPUT OUTPUT [-a estimate|info] [-f GT|GL|PL] [-t THREADS]
//           [-i ITERATIONS] [-m MAF] [-s SAMPLES]
//
// -f GT (default) runs the CLS / Brier estimator (coefficients.rs:166-278);
// -f GL / PL run the maximum-likelihood fit (coefficients.rs:40-141): on the
// 2-bit cohort when every likelihood row is one-hot (hard calls), else on the
// likelihood rows.
#include <sys/resource.h>

#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/kestrel_amd.h"
#include "vcf.hpp"

namespace {

const char* VERSION = "0.1.2";

void info(const std::string& m) { fprintf(stderr, "[INFO ] %s\n", m.c_str()); }
void warn(const std::string& m) { fprintf(stderr, "[WARN ] %s\n", m.c_str()); }

[[noreturn]] void usage_error(const std::string& m) {
    fprintf(stderr, "error: %s\n\nUsage: kestrel [OPTIONS] <INPUT> <OUTPUT>\n\nFor more information, try '--help'.\n",
            m.c_str());
    exit(2);
}

void help() {
    printf(
        "Pairwise Jacquard coefficients and kinship from a VCF, on an MI355X GPU\n\n"
        "Usage: kestrel [OPTIONS] <INPUT> <OUTPUT>\n\n"
        "Arguments:\n"
        "  <INPUT>   VCF file with genotypes (GT) or genotype likelihoods (GL, PL)\n"
        "  <OUTPUT>  Tab-separated file to write the coefficients to\n\n"
        "Options:\n"
        "  -a, --allele-freqs <ALLELE_FREQS>  Where the allele frequencies come from [default: estimate]\n"
        "                                     [possible values: estimate, info]\n"
        "  -f, --field <FIELD>                Which FORMAT field to compute kinship from [default: GT]\n"
        "                                     [possible values: GT, GL, PL]\n"
        "  -t, --threads <THREADS>            Number of threads reading the VCF (the fit runs on the GPU)\n"
        "  -i, --iterations <ITERATIONS>      Maximum number of iterations for minimization algorithm [default: 100]\n"
        "  -m, --minimum-allele-frequency <MINIMUM_ALLELE_FREQUENCY>\n"
        "                                     Allele frequencies below this threshold are set to zero\n"
        "  -s, --samples <SAMPLES>            Only calculate the Jacquard coefficients for the first n samples\n"
        "  -h, --help                         Print help\n"
        "  -V, --version                      Print version\n");
}

// ryu's shortest round-trip text (the csv crate writes f64 with ryu):
// shortest digits, then decimal for exponents in [-5, 16], scientific
// otherwise; "NaN", "inf", "-inf".
std::string fmt_f64(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
    std::string s(buf, r.ptr);
    std::string out;
    if (s[0] == '-') {
        out = "-";
        s = s.substr(1);
    }
    const size_t epos = s.find('e');
    const int exp10 = atoi(s.c_str() + epos + 1);
    std::string digits;
    for (size_t i = 0; i < epos; i++)
        if (s[i] != '.') digits += s[i];
    const int len = (int)digits.size();
    const int e = exp10 - (len - 1);  // value = digits * 10^e
    const int kk = len + e;           // position of the decimal point
    if (0 <= e && kk <= 16) {
        out += digits + std::string(e, '0') + ".0";
    } else if (0 < kk && kk <= 16) {
        out += digits.substr(0, kk) + "." + digits.substr(kk);
    } else if (-5 < kk && kk <= 0) {
        out += "0." + std::string(-kk, '0') + digits;
    } else if (len == 1) {
        out += digits + "e" + std::to_string(kk - 1);
    } else {
        out += digits.substr(0, 1) + "." + digits.substr(1) + "e" + std::to_string(kk - 1);
    }
    return out;
}

std::string tsv_field(const std::string& s) {
    // csv quoting for fields holding the delimiter, quotes or line breaks
    if (s.find_first_of("\t\"\n\r") == std::string::npos) return s;
    std::string q = "\"";
    for (char ch : s) q += ch == '"' ? std::string("\"\"") : std::string(1, ch);
    return q + "\"";
}

std::string duration(double s) {
    char b[64];
    if (s >= 1.0) snprintf(b, sizeof b, "%.2fs", s);
    else snprintf(b, sizeof b, "%.2fms", s * 1e3);
    return b;
}

// util.rs:9-38 ResourceTracker line (parsed by analyze-real.py:12-31)
struct Tracker {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    rusage u0{};
    Tracker() { getrusage(RUSAGE_SELF, &u0); }
    void checkpoint() {
        rusage u{};
        getrusage(RUSAGE_SELF, &u);
        auto tv = [](const timeval& a, const timeval& b) {
            return (double)(b.tv_sec - a.tv_sec) + (double)(b.tv_usec - a.tv_usec) * 1e-6;
        };
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        char b[256];
        snprintf(b, sizeof b, "Wall Time %s, User Time %s, Kernel Time %s, Peak Memory %.2f MiB", duration(wall).c_str(),
                 duration(tv(u0.ru_utime, u.ru_utime)).c_str(), duration(tv(u0.ru_stime, u.ru_stime)).c_str(),
                 (double)u.ru_maxrss / 1024.0);
        info(b);
    }
};

struct Args {
    std::string input, output;
    bool af_info = false;
    kvcf::Field field = kvcf::Field::GT;
    long threads = 0;
    uint64_t iterations = 100;
    double maf = -1.0;
    bool have_maf = false;
    long samples = -1;
    bool dump_codes = false;     // print the decoded calls and stop (no GPU)
    std::string dump_probs;      // write the GL / PL likelihood rows to this file and stop (no GPU)
    bool format_floats = false;  // print stdin doubles as TSV text and stop
};

Args parse(int argc, char** argv) {
    Args a;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; i++) {
        std::string s = argv[i];
        std::string val;
        auto take = [&](const char* sh, const char* lo) -> bool {
            const size_t ls = strlen(lo);
            if (s == sh || s == lo) {
                if (i + 1 >= argc) usage_error(std::string("a value is required for '") + lo + "'");
                val = argv[++i];
                return true;
            }
            if (s.compare(0, ls + 1, std::string(lo) + "=") == 0) {
                val = s.substr(ls + 1);
                return true;
            }
            if (s.size() > 2 && s.compare(0, 2, sh) == 0 && s[1] != '-') {
                val = s.substr(2);
                return true;
            }
            return false;
        };
        if (s == "--dump-codes") {
            a.dump_codes = true;
            continue;
        }
        if (s == "--dump-probs" && i + 1 < argc) {  // GL / PL likelihood rows as raw doubles
            a.dump_probs = argv[++i];
            continue;
        }
        if (s == "--format-floats") {
            a.format_floats = true;
            return a;
        }
        if (s == "-h" || s == "--help") {
            help();
            exit(0);
        } else if (s == "-V" || s == "--version") {
            printf("kestrel %s (kestrel_amd, gfx950)\n", VERSION);
            exit(0);
        } else if (take("-a", "--allele-freqs")) {
            if (val == "estimate") a.af_info = false;
            else if (val == "info") a.af_info = true;
            else usage_error("invalid value '" + val + "' for '--allele-freqs <ALLELE_FREQS>'");
        } else if (take("-f", "--field")) {
            if (val == "GT") a.field = kvcf::Field::GT;
            else if (val == "GL") a.field = kvcf::Field::GL;
            else if (val == "PL") a.field = kvcf::Field::PL;
            else usage_error("invalid value '" + val + "' for '--field <FIELD>'");
        } else if (take("-t", "--threads")) {
            char* e = nullptr;
            a.threads = strtol(val.c_str(), &e, 10);
            if (*e || a.threads <= 0) usage_error("invalid value '" + val + "' for '--threads <THREADS>'");
        } else if (take("-i", "--iterations")) {
            char* e = nullptr;
            long long it = strtoll(val.c_str(), &e, 10);
            if (*e || it <= 0) usage_error("invalid value '" + val + "' for '--iterations <ITERATIONS>'");
            a.iterations = (uint64_t)it;
        } else if (take("-m", "--minimum-allele-frequency")) {
            char* e = nullptr;
            a.maf = strtod(val.c_str(), &e);
            if (*e) usage_error("invalid value '" + val + "' for '--minimum-allele-frequency'");
            a.have_maf = true;
        } else if (take("-s", "--samples")) {
            char* e = nullptr;
            a.samples = strtol(val.c_str(), &e, 10);
            if (*e || a.samples < 0) usage_error("invalid value '" + val + "' for '--samples <SAMPLES>'");
        } else if (!s.empty() && s[0] == '-' && s != "-") {
            usage_error("unexpected argument '" + s + "' found");
        } else {
            pos.push_back(s);
        }
    }
    if (pos.size() < 2) usage_error("the following required arguments were not provided: <INPUT> <OUTPUT>");
    if (pos.size() > 2) usage_error("unexpected argument '" + pos[2] + "' found");
    a.input = pos[0];
    a.output = pos[1];
    return a;
}

int die(const std::string& m) {
    fprintf(stderr, "Error: %s\n", m.c_str());
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    Args a = parse(argc, argv);
    if (a.format_floats) {
        char line[256];
        while (fgets(line, sizeof line, stdin)) printf("%s\n", fmt_f64(strtod(line, nullptr)).c_str());
        return 0;
    }
    info(std::string("Kestrel ") + VERSION + " (" + kestrel_version() + ")");
    info("Parsing VCF \"" + a.input + "\"");
    const auto t_start = std::chrono::steady_clock::now();
    kvcf::Data d;
    std::string err;
    if (!kvcf::read(a.input, a.field, a.af_info, d, err, a.threads > 0 ? (unsigned)a.threads : 0)) return die(err);
    const auto t_parsed = std::chrono::steady_clock::now();
    info("Parsed " + std::to_string(d.total) + " total sites");
    info("Skipped " + std::to_string(d.not_snp) + " sites that were not SNPs");
    if (a.field == kvcf::Field::GT) info("Skipped " + std::to_string(d.skipped_missing) + " sites with missing data");
    info("Found " + std::to_string(d.n_sites) + " di-allelic sites");
    info("Found " + std::to_string(d.n_sites3) + " tri-allelic sites");
    info("Found " + std::to_string(d.n_sites4) + " tetra-allelic sites");
    const uint32_t n = (uint32_t)d.samples.size();
    if (!a.dump_probs.empty()) {
        // the 2-, 3- and 4-allele buckets in turn as raw doubles, site-major
        // (n_sites x n_samples x 3, then x 6, then x 10: the rows the
        // likelihood paths receive)
        FILE* f = fopen(a.dump_probs.c_str(), "wb");
        bool ok = f != nullptr;
        d.join_probs();
        for (const std::vector<double>* v : {&d.probs, &d.probs3, &d.probs4})
            ok = ok && (v->empty() || fwrite(v->data(), sizeof(double), v->size(), f) == v->size());
        if (f) fclose(f);
        if (!ok) return die("cannot write " + a.dump_probs);
        return 0;
    }
    if (a.dump_codes) {
        // samples, then one line per kept site: codes (0/1/2, 3 = missing) and AF
        for (uint32_t i = 0; i < n; i++) printf("%s%s", i ? "\t" : "", d.samples[i].c_str());
        printf("\n");
        std::string line;
        for (uint32_t s = 0; s < d.n_sites; s++) {
            line.assign(n, '0');
            for (uint32_t i = 0; i < n; i++) line[i] = (char)('0' + d.site_codes[(size_t)s * n + i]);
            if (a.af_info) line += "\t" + fmt_f64(d.af[2 * s]) + "\t" + fmt_f64(d.af[2 * s + 1]);
            line += '\n';
            fwrite(line.data(), 1, line.size(), stdout);
        }
        // GT sites with 3 and 4 alleles: "A:" then their genotype indices (and AF)
        for (int A = 3; A <= 4; A++) {
            const uint32_t v = A == 3 ? d.n_sites3 : d.n_sites4;
            const std::vector<uint8_t>& codes = A == 3 ? d.site_codes3 : d.site_codes4;
            const std::vector<double>& afs = A == 3 ? d.af3 : d.af4;
            if (a.field != kvcf::Field::GT) break;
            for (uint32_t s = 0; s < v; s++) {
                line = std::to_string(A) + ":";
                for (uint32_t i = 0; i < n; i++) line += (char)('0' + codes[(size_t)s * n + i]);
                if (a.af_info)
                    for (int k = 0; k < A; k++) line += "\t" + fmt_f64(afs[(size_t)s * A + k]);
                line += '\n';
                fwrite(line.data(), 1, line.size(), stdout);
            }
        }
        fprintf(stderr, "total=%llu not_snp=%llu skipped_missing=%llu\n", (unsigned long long)d.total,
                (unsigned long long)d.not_snp, (unsigned long long)d.skipped_missing);
        return 0;
    }
    if ((uint64_t)d.n_sites + d.n_sites3 + d.n_sites4 == 0) return die("Must have at least one site");
    if (n == 0) return die("Must have at least one sample");
    const uint32_t nsel = a.samples < 0 ? n : (uint32_t)a.samples;
    if (nsel > n) {
        fprintf(stderr, "Requested samples %u exceeds total samples %u\n", nsel, n);
        return 1;
    }
    if (a.threads) info("Read the VCF with " + std::to_string(a.threads) + " threads; the fit runs on the GPU");
    size_t row_bytes = 0;
    std::vector<uint8_t> packed = d.pack(&row_bytes);
    const double* af = a.af_info ? d.af.data() : nullptr;

    Tracker tracker;
    kestrel_cohort* c = nullptr;
    const size_t npairs = (size_t)nsel * (nsel + 1) / 2;
    std::vector<kestrel_output> out(npairs);
    int rc;
    if (a.field == kvcf::Field::GT) {
        if (d.n_sites3 || d.n_sites4) {  // coefficients.rs:166-320 over the three buckets
            const std::vector<uint8_t> gt3 = d.sample_major(3), gt4 = d.sample_major(4);
            rc = kestrel_cohort_create_gt_multi(&c, n, d.n_sites ? packed.data() : nullptr, d.n_sites, row_bytes,
                                                d.n_sites3 ? gt3.data() : nullptr, d.n_sites3,
                                                d.n_sites4 ? gt4.data() : nullptr, d.n_sites4,
                                                a.af_info ? d.af.data() : nullptr, a.af_info ? d.af3.data() : nullptr,
                                                a.af_info ? d.af4.data() : nullptr);
        } else {
            rc = kestrel_cohort_create_gt(&c, packed.data(), n, d.n_sites, row_bytes, af);
        }
        if (rc) return die(kestrel_last_error());
        info("Calculating Jacquard coefficients for " + std::to_string(npairs) + " pairs using " +
             std::to_string((uint64_t)d.n_sites + d.n_sites3 + d.n_sites4) + " sites");
        rc = kestrel_fit_pairs_gt(c, a.iterations, nullptr, 0, nsel, out.data(), nullptr);
        if (rc) return die(kestrel_last_error());
        for (const auto& o : out)
            if (o.iters >= a.iterations)
                warn("No convergence for Jacquard coefficients, max iterations " + std::to_string(a.iterations) +
                     " exceeded");
    } else {
        const double maf = a.have_maf ? a.maf : KESTREL_MAF_DEFAULT;
        if (a.have_maf && !(maf >= 0.0 && maf <= 1.0)) {
            fprintf(stderr, "Minimum allele frequency must be within 0 to 1\n");
            return 1;
        }
        info("Calculating allele frequencies for " + std::to_string(d.n_sites) + " sites using " + std::to_string(n) +
             " samples");
        // hard calls take the 2-bit path; any soft row, the likelihood path;
        // sites with 3 or 4 alleles, the multi-allelic likelihood path
        const bool multi = d.n_sites3 || d.n_sites4;
        if (multi || !d.one_hot) d.join_probs();
        if (multi)
            rc = kestrel_cohort_create_gl_multi(&c, n, d.n_sites ? d.probs.data() : nullptr, d.n_sites,
                                                d.n_sites3 ? d.probs3.data() : nullptr, d.n_sites3,
                                                d.n_sites4 ? d.probs4.data() : nullptr, d.n_sites4, maf, a.iterations);
        else if (d.one_hot)
            rc = kestrel_cohort_create(&c, packed.data(), n, d.n_sites, row_bytes, af, maf, a.iterations,
                                       KESTREL_FP64);
        else
            rc = kestrel_cohort_create_gl(&c, d.probs.data(), n, d.n_sites, af, maf, a.iterations);
        if (rc) return die(kestrel_last_error());
        tracker.checkpoint();
        uint32_t parts[3] = {0, 0, 0};
        if (multi)
            kestrel_cohort_multi_info(c, parts, nullptr, nullptr, nullptr);
        else
            kestrel_cohort_n_kept(c, &parts[0]);
        const uint64_t input = (uint64_t)d.n_sites + d.n_sites3 + d.n_sites4;
        const uint32_t kept = parts[0] + parts[1] + parts[2];
        info("Identified " + std::to_string(parts[0]) + " di-allelic sites");
        info("Identified " + std::to_string(parts[1]) + " tri-allelic sites");
        info("Identified " + std::to_string(parts[2]) + " tetra-allelic sites");
        info("Discarded " + std::to_string(input - kept) + " non-segregating sites");
        info("Calculating Jacquard coefficients for " + std::to_string(npairs) + " pairs using " +
             std::to_string(kept) + " sites");
        kestrel_tune t;
        kestrel_tune_default(&t, a.iterations);
        rc = kestrel_fit_pairs(c, &t, nullptr, 0, nsel, out.data());
        if (rc) return die(kestrel_last_error());
        for (const auto& o : out) {
            if (o.status == KESTREL_PAIR_NO_SITES)
                warn("pair " + std::to_string(o.x) + " " + std::to_string(o.y) + " has no usable site");
            else if (o.iters >= a.iterations)
                warn("No convergence for Jacquard coefficients of pair " + std::to_string(o.x) + " " +
                     std::to_string(o.y) + ", max iterations " + std::to_string(a.iterations) + " exceeded");
        }
    }
    tracker.checkpoint();
    kestrel_cohort_destroy(c);
    const auto t_fitted = std::chrono::steady_clock::now();

    // cli.rs:129-181
    FILE* f = fopen(a.output.c_str(), "w");
    if (!f) return die("No such file or directory (os error 2): " + a.output);
    fputs("sample1\tsample2\tkinship\tjacquard1\tjacquard2\tjacquard3\tjacquard4\tjacquard5\tjacquard6\tjacquard7\t"
          "jacquard8\tjacquard9\titerations\tobjective\tcondition\n",
          f);
    static const double KIN[9] = {1.0, 0.0, 0.5, 0.0, 0.5, 0.0, 0.5, 0.25, 0.0};
    auto format_rows = [&](size_t lo, size_t hi, std::string& buf) {
        for (size_t i = lo; i < hi; i++) {
            const kestrel_output& o = out[i];
            double kin = 0.0;
            for (int k = 0; k < 9; k++) kin = std::fma(o.jacquard[k], KIN[k], kin);  // algebra.rs dot
            buf += tsv_field(d.samples[o.x]);
            buf += '\t';
            buf += tsv_field(d.samples[o.y]);
            buf += '\t';
            buf += fmt_f64(kin);
            for (int k = 0; k < 9; k++) {
                buf += '\t';
                buf += fmt_f64(o.jacquard[k]);
            }
            buf += '\t';
            buf += std::to_string(o.iters);
            buf += '\t';
            buf += fmt_f64(o.obj);
            buf += '\t';
            buf += fmt_f64(std::log10(o.cond));
            buf += '\n';
        }
    };
    // rows are formatted in parallel chunks and written in order
    const size_t nthreads = std::max<size_t>(1, a.threads > 0 ? (size_t)a.threads
                                                              : (size_t)std::thread::hardware_concurrency());
    const size_t chunk = 16384;
    for (size_t base = 0; base < out.size(); base += chunk * nthreads) {
        std::vector<std::string> bufs(nthreads);
        std::vector<std::thread> pool;
        for (size_t t = 0; t < nthreads; t++) {
            const size_t lo = std::min(out.size(), base + t * chunk), hi = std::min(out.size(), lo + chunk);
            if (lo < hi) pool.emplace_back(format_rows, lo, hi, std::ref(bufs[t]));
        }
        for (auto& th : pool) th.join();
        for (const auto& b : bufs)
            if (!b.empty() && fwrite(b.data(), 1, b.size(), f) != b.size()) {
                fclose(f);
                return die("failed to write " + a.output);
            }
    }
    if (fclose(f) != 0) return die("failed to write " + a.output);
    auto secs = [](std::chrono::steady_clock::time_point x, std::chrono::steady_clock::time_point y) {
        char b[32];
        snprintf(b, sizeof b, "%.3f s", std::chrono::duration<double>(y - x).count());
        return std::string(b);
    };
    info("Phases: parse " + secs(t_start, t_parsed) + ", device setup and fit " + secs(t_parsed, t_fitted) +
         ", write " + secs(t_fitted, std::chrono::steady_clock::now()));
    return 0;
}
