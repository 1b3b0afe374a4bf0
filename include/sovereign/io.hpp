// Model readers.  API mirrors include/sovereign/io.hpp:24-67 of the
// reference.  This project keeps a compact CPLEX-LP reader (the subset the
// engine's tests use) and a free/fixed MPS reader; gzip input is not
// supported (no zlib dependency) and is rejected with a clear error.
#pragma once

#include <string>
#include <vector>

#include "sovereign/problem.hpp"

namespace sov {

struct ReaderWarning {
    int line = 0;
    std::string section;
    std::string message;
};

struct ReadResult {
    Problem problem;
    std::vector<ReaderWarning> warnings;
    std::string format;
};

enum class NegativeUpperBoundRule { LowerBecomesMinusInfinity, KeepLowerAtZero };

struct MpsOptions {
    NegativeUpperBoundRule negative_upper_bound = NegativeUpperBoundRule::LowerBecomesMinusInfinity;
    bool keep_free_rows = true;
};

ReadResult read_mps(const std::string& path, const MpsOptions& options = {});
ReadResult read_mps_string(const std::string& text, const MpsOptions& options = {},
                           const std::string& origin = "<string>");
ReadResult read_lp(const std::string& path);
ReadResult read_lp_string(const std::string& text, const std::string& origin = "<string>");
ReadResult read_model(const std::string& path);

/// Additive: the model as free-format MPS text that read_mps reads back to
/// the same arrays bit for bit (values in shortest round-trip decimal form).
/// Rows and columns take the model's names (R<i> / C<j> for a model built
/// without them); two-sided rows are written as G rows with a range, the
/// objective constant as an RHS entry on the objective row.
std::string write_mps_string(const Problem& problem);
void write_mps(const Problem& problem, const std::string& path);

}  // namespace sov
