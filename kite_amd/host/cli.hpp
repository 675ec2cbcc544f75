// cli.hpp -- command-line parsing shaped like the reference's clap
// definitions (src/cli.rs): long and short flags, `--flag value` and
// `--flag=value`, and multi-value flags that take arguments until the next
// flag.
#pragma once

#include <cstdint>
#include <map>
#include <optional>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace kh {

struct UsageError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct FlagSpec {
    std::string name;       // long name without dashes
    char short_name = 0;    // 0: none
    bool takes_value = true;
    bool multi = false;     // one or more values
};

class Args {
   public:
    Args(const std::vector<FlagSpec> &specs, int argc, char **argv, int first);
    bool has(const std::string &n) const { return vals_.count(n) > 0; }
    std::string get(const std::string &n, const std::string &def = "") const;
    std::vector<std::string> list(const std::string &n) const;
    std::optional<uint32_t> get_u32(const std::string &n) const;
    uint64_t get_u64(const std::string &n, uint64_t def) const;
    bool get_bool(const std::string &n, bool def) const;
    bool flag(const std::string &n) const { return has(n); }

   private:
    std::map<std::string, std::vector<std::string>> vals_;
};

uint64_t parse_u64(const std::string &s, const std::string &what);

}  // namespace kh
