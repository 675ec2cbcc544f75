// cli.cpp -- flag parsing (see cli.hpp).
#include "cli.hpp"

#include <cerrno>
#include <cstdlib>

namespace kh {

uint64_t parse_u64(const std::string &s, const std::string &what) {
    if (s.empty()) throw UsageError("invalid value '' for '--" + what + "'");
    for (char c : s)
        if (c < '0' || c > '9') throw UsageError("invalid value '" + s + "' for '--" + what + "': invalid digit found in string");
    errno = 0;
    unsigned long long v = strtoull(s.c_str(), nullptr, 10);
    if (errno) throw UsageError("invalid value '" + s + "' for '--" + what + "': number too large");
    return v;
}

Args::Args(const std::vector<FlagSpec> &specs, int argc, char **argv, int first) {
    auto find_long = [&](const std::string &n) -> const FlagSpec * {
        for (auto &s : specs)
            if (s.name == n) return &s;
        return nullptr;
    };
    auto find_short = [&](char c) -> const FlagSpec * {
        for (auto &s : specs)
            if (s.short_name == c) return &s;
        return nullptr;
    };
    for (int i = first; i < argc; i++) {
        std::string a = argv[i];
        const FlagSpec *sp = nullptr;
        std::string inline_val;
        bool has_inline = false;
        if (a.rfind("--", 0) == 0) {
            std::string n = a.substr(2);
            size_t eq = n.find('=');
            if (eq != std::string::npos) {
                inline_val = n.substr(eq + 1);
                n = n.substr(0, eq);
                has_inline = true;
            }
            sp = find_long(n);
            if (!sp) throw UsageError("unexpected argument '" + a + "' found");
        } else if (a.size() >= 2 && a[0] == '-' && a[1] != '-') {
            sp = find_short(a[1]);
            if (!sp) throw UsageError("unexpected argument '" + a + "' found");
            if (a.size() > 2) {
                inline_val = a.substr(2);
                if (inline_val[0] == '=') inline_val = inline_val.substr(1);
                has_inline = true;
            }
        } else {
            throw UsageError("unexpected argument '" + a + "' found");
        }
        auto &slot = vals_[sp->name];
        if (!sp->takes_value) {
            if (has_inline) throw UsageError("unexpected value for '--" + sp->name + "'");
            slot.push_back("true");
            continue;
        }
        if (!sp->multi) slot.clear();
        if (has_inline) {
            slot.push_back(inline_val);
        } else {
            if (i + 1 >= argc) throw UsageError("a value is required for '--" + sp->name + "' but none was supplied");
            slot.push_back(argv[++i]);
        }
        if (sp->multi) {
            while (i + 1 < argc && !(argv[i + 1][0] == '-' && argv[i + 1][1] != 0)) slot.push_back(argv[++i]);
        }
    }
}

std::string Args::get(const std::string &n, const std::string &def) const {
    auto it = vals_.find(n);
    return it == vals_.end() || it->second.empty() ? def : it->second.back();
}

std::vector<std::string> Args::list(const std::string &n) const {
    auto it = vals_.find(n);
    return it == vals_.end() ? std::vector<std::string>{} : it->second;
}

std::optional<uint32_t> Args::get_u32(const std::string &n) const {
    if (!has(n)) return std::nullopt;
    uint64_t v = parse_u64(get(n), n);
    if (v > 0xFFFFFFFFull) throw UsageError("invalid value '" + get(n) + "' for '--" + n + "': number too large to fit in target type");
    return (uint32_t)v;
}

uint64_t Args::get_u64(const std::string &n, uint64_t def) const {
    if (!has(n)) return def;
    return parse_u64(get(n), n);
}

bool Args::get_bool(const std::string &n, bool def) const {
    if (!has(n)) return def;
    std::string v = get(n);
    if (v == "true") return true;
    if (v == "false") return false;
    throw UsageError("invalid value '" + v + "' for '--" + n + " <BOOL>'\n  [possible values: true, false]");
}

}  // namespace kh
