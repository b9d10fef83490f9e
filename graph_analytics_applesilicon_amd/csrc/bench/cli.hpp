// Small helpers shared by the command-line tools (gbench, gconvert, gstats).
#pragma once

#include <sys/stat.h>

#include <fstream>
#include <string>
#include <vector>

namespace cli {

inline bool file_exists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

// File name without directories and without everything from its first '.'
// ("data/lj.txt.gz" -> "lj"), the dataset key of the CSV files.
inline std::string stem(const std::string& path) {
    const size_t s = path.find_last_of('/');
    const std::string base = s == std::string::npos ? path : path.substr(s + 1);
    return base.substr(0, base.find('.'));
}

// Rewrites `path` with `header`, every old row whose first field is not `key`,
// then `rows`: re-running a tool on a dataset replaces its rows.
inline void csv_upsert(const std::string& path, const std::string& header, const std::string& key,
                       const std::vector<std::string>& rows) {
    std::vector<std::string> keep;
    {
        std::ifstream in(path);
        std::string line;
        for (bool first = true; std::getline(in, line); first = false)
            if (!first && !line.empty() && line.substr(0, line.find(',')) != key) keep.push_back(line);
    }
    std::ofstream out(path, std::ios::trunc);
    out << header << '\n';
    for (const auto& r : keep) out << r << '\n';
    for (const auto& r : rows) out << r << '\n';
}

}  // namespace cli
