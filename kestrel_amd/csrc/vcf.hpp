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
    uint64_t total = 0, not_snp = 0, skipped_missing = 0, multiallelic = 0;

    // 2-bit rows, sample-major, site s of a row in bits 2*(s%4) of byte s/4
    std::vector<uint8_t> pack(size_t* row_bytes) const;
};

// Reads `path`; returns false with `err` set on failure (the reference's
// anyhow errors).  GT: a site with a missing call is skipped (vcf.rs:217-221).
// GL / PL: likelihood rows become codes when they are one-hot (the called
// genotype has probability 1, the others 0) and code 3 when a value is
// missing (vcf.rs:355, 397).
bool read(const std::string& path, Field field, bool parse_af, Data& out, std::string& err);

}  // namespace kvcf
