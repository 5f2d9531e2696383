// vcf.hpp -- VCF reader of the kestrel CLI: plain or gzip/bgzf text to 2-bit
// hard-call codes (vcf.rs:96-482).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace kvcf {

enum class Field { GT, GL, PL };

struct Data {
    std::vector<std::string> samples;
    uint32_t n_sites = 0;
    std::vector<uint8_t> site_codes;  // n_sites x n_samples (site-major), 0/1/2 genotype index, 3 missing
    std::vector<double> af;           // n_sites x 2 when INFO AF was requested
    std::vector<double> probs;        // GL / PL: n_sites x n_samples x 3 linear-scale likelihoods (NaN = missing),
                                      // filled by join_probs()
    std::vector<std::vector<double>> probs_parts;  // the same rows as read, batch by batch
    bool one_hot = true;              // GL / PL: every called row has one 1 and two 0
    // GL / PL sites with 3 and 4 alleles (vcf.rs:270-305 buckets), site-major
    // n_sites3 x n_samples x 6 and n_sites4 x n_samples x 10
    uint32_t n_sites3 = 0, n_sites4 = 0;
    std::vector<double> probs3, probs4;
    // GT sites with 3 and 4 alleles: site-major genotype indices (VCF order) and
    // INFO AF (n_sites3 x 3, n_sites4 x 4; reference allele first)
    std::vector<uint8_t> site_codes3, site_codes4;
    std::vector<double> af3, af4;
    uint64_t total = 0, not_snp = 0, skipped_missing = 0, multiallelic = 0;

    // 2-bit rows, sample-major, site s of a row in bits 2*(s%4) of byte s/4
    std::vector<uint8_t> pack(size_t* row_bytes) const;
    // Moves probs_parts into probs (one copy; the 2-bit path of one-hot rows
    // never needs it)
    void join_probs();
    // The GT genotype indices of the 3- or 4-allele sites, sample-major (N x V_A)
    std::vector<uint8_t> sample_major(int alleles) const;
};

// Reads `path`; returns false with `err` set on failure (the reference's
// anyhow errors).  GT: a site with a missing call is skipped (vcf.rs:217-221).
// GL / PL: every row is kept as its linear-scale likelihoods (vcf.rs:368, 407)
// in `probs` (sites with 3 or 4 alleles in `probs3` / `probs4`; GT sites with 3
// or 4 alleles in `site_codes3` / `site_codes4`); a row with a NaN (a missing value, vcf.rs:355, 397) has code 3,
// a one-hot row (the called genotype 1, the others 0) the called genotype's
// code; `one_hot` says whether every called row was one-hot.
// threads: parser / inflater threads (0 = one per hardware thread); the
// result does not depend on it.
bool read(const std::string& path, Field field, bool parse_af, Data& out, std::string& err, unsigned threads = 0);

}  // namespace kvcf
