// Header-only stand-in for the reference's Config (src/config.h:50-82,
// src/config.cpp): an INI file of [section] blocks with "key = value" lines,
// overridden from the command line by --section:key=value, read through
// typed get(section, key, default).  Same class name, member functions and
// parsing rules, so a caller written against the reference (NSCube's
// constructor, test/test_ns_cube.cpp's main) compiles against it unchanged:
//
//   * a line that starts with '[' names a section: the first whitespace-free
//     word after the bracket (leading blanks skipped), less its last
//     character -- "[ns]" is "ns", "[ns ]" is "n";
//   * lines that start with '#' or ';' are comments, and ';' ends a line;
//   * otherwise, once a section is open, the first two words separated by
//     blanks, '=' or tabs are key and value (further words are ignored);
//     lines before the first section are ignored;
//   * a missing file leaves the configuration empty (every get returns its
//     default).
// Values are converted as the reference's lexical_cast does (stream
// extraction: "1.5" read as an int is 1); a value that does not convert at
// all reads as 0, where the reference leaves the result uninitialised.
#pragma once

#include <cstdio>
#include <cstring>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

struct ConfigSkeleton {
    enum Flags { OPTIONAL = 0, REQUIRED = 1 };
    // parameter name -> (flags, docstring)
    typedef std::map<std::string, std::pair<Flags, std::string>> section_t;
    // section name -> section
    typedef std::map<std::string, section_t> data_t;

    data_t data;

    bool is_required(const std::string& section, const std::string& prm) const {
        auto s = data.find(section);
        if (s == data.end()) return false;
        auto p = s->second.find(prm);
        return p != s->second.end() && (p->second.first & REQUIRED);
    }
};

class Config {
    typedef std::map<std::string, std::string> section_t;
    typedef std::map<std::string, section_t> data_t;

    data_t data;
    std::string filename;
    ConfigSkeleton skeleton;

    template <typename A>
    static A convert(const std::string& s) {
        A a{};
        std::istringstream in(s);
        in >> a;
        return a;
    }

public:
    void set_skeleton(const ConfigSkeleton& s) { skeleton = s; }

    void open(const std::string& fn) {
        filename = fn;
        FILE* f = std::fopen(fn.c_str(), "rb");
        if (!f) return;
        load(f);
        std::fclose(f);
    }

    void load(FILE* f) {
        char line[4096];
        section_t* cur = nullptr;
        while (std::fgets(line, sizeof line, f)) {
            std::string s(line);
            if (!s.empty() && s[0] == '[') {
                const size_t b = s.find_first_not_of(" \t\n\r\v\f", 1);
                if (b != std::string::npos) {
                    const size_t e = s.find_first_of(" \t\n\r\v\f", b);
                    std::string word = s.substr(b, (e == std::string::npos ? s.size() : e) - b);
                    word.pop_back();
                    cur = &data[word];
                    continue;
                }
            }
            if (!s.empty() && (s[0] == '#' || s[0] == ';')) continue;
            if (!cur) continue;
            if (!s.empty() && s.back() == '\n') s.pop_back();
            if (!s.empty() && s.back() == '\r') s.pop_back();
            const size_t semi = s.find(';');
            if (semi != std::string::npos) s.resize(semi);
            const char* sep = " =\t";
            const size_t k0 = s.find_first_not_of(sep);
            if (k0 == std::string::npos) continue;
            const size_t k1 = s.find_first_of(sep, k0);
            if (k1 == std::string::npos) continue;
            const size_t v0 = s.find_first_not_of(sep, k1);
            if (v0 == std::string::npos) continue;
            const size_t v1 = s.find_first_of(sep, v0);
            (*cur)[s.substr(k0, k1 - k0)] = s.substr(v0, v1 == std::string::npos ? std::string::npos : v1 - v0);
        }
    }

    // --section_name:option_name=value; any other argument is skipped
    void rewrite(int argc, char** argv) {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            if (a.compare(0, 2, "--") != 0) continue;
            const size_t colon = a.find(':', 2);
            if (colon == std::string::npos) continue;
            const size_t eq = a.find('=', colon + 1);
            if (eq == std::string::npos) continue;
            data[a.substr(2, colon - 2)][a.substr(colon + 1, eq - colon - 1)] = a.substr(eq + 1);
        }
    }

    // false (and the missing key on stderr) when a REQUIRED parameter of a
    // section that is present is absent
    bool validate() const {
        for (const auto& sec : skeleton.data) {
            auto have = data.find(sec.first);
            if (have == data.end()) continue;
            for (const auto& prm : sec.second)
                if (have->second.find(prm.first) == have->second.end() &&
                    (prm.second.first & ConfigSkeleton::REQUIRED)) {
                    std::fprintf(stderr, "%s:%s REQUIRED\n", sec.first.c_str(), prm.first.c_str());
                    return false;
                }
        }
        return true;
    }

    void help() const {
        for (const auto& sec : skeleton.data)
            for (const auto& prm : sec.second)
                std::fprintf(stderr, "--%s:%s -- %s -- %s\n", sec.first.c_str(), prm.first.c_str(),
                             (prm.second.first & ConfigSkeleton::REQUIRED) ? "REQUIRED" : "OPTIONAL",
                             prm.second.second.c_str());
    }

    void save() const {
        FILE* f = std::fopen(filename.c_str(), "wb");
        if (!f) return;
        print(f);
        std::fclose(f);
    }

    void print(FILE* f) const {
        for (const auto& sec : data) {
            std::fprintf(f, "[%s]\n", sec.first.c_str());
            for (const auto& kv : sec.second) std::fprintf(f, "%s = %s\n", kv.first.c_str(), kv.second.c_str());
            std::fprintf(f, "\n");
        }
    }

    void print(std::string& str) const {
        std::string out;
        for (const auto& sec : data) {
            out += "[" + sec.first + "]\n";
            for (const auto& kv : sec.second) out += kv.first + " = " + kv.second + "\n";
            out += "\n";
        }
        str = out;
    }

    // With a default: the default when the section or key is absent; a
    // parameter the skeleton marks REQUIRED cannot be read with a default.
    std::string get(const std::string& section, const std::string& name, const std::string& def) const {
        if (skeleton.is_required(section, name))
            throw std::runtime_error("cannot use with default\nparameter " + section + ":" + name + " is required");
        auto s = data.find(section);
        if (s == data.end()) return def;
        auto k = s->second.find(name);
        return k == s->second.end() ? def : k->second;
    }
    std::string get(const std::string& section, const std::string& name, const char* def) const {
        return get(section, name, std::string(def));
    }
    double get(const std::string& section, const std::string& name, double def) const {
        if (skeleton.is_required(section, name) || has(section, name))
            return convert<double>(get(section, name, std::string()));
        return def;
    }
    int get(const std::string& section, const std::string& name, int def) const {
        if (skeleton.is_required(section, name) || has(section, name))
            return convert<int>(get(section, name, std::string()));
        return def;
    }

    // Without a default: std::runtime_error when the section or key is absent.
    std::string gets(const std::string& section, const std::string& name) const {
        auto s = data.find(section);
        if (s == data.end()) throw std::runtime_error("section " + section + " not found");
        auto k = s->second.find(name);
        if (k == s->second.end()) throw std::runtime_error("parameter " + section + ":" + name + " not found");
        return k->second;
    }
    double getd(const std::string& section, const std::string& name) const {
        return convert<double>(gets(section, name));
    }
    int geti(const std::string& section, const std::string& name) const { return convert<int>(gets(section, name)); }

private:
    bool has(const std::string& section, const std::string& name) const {
        auto s = data.find(section);
        return s != data.end() && s->second.count(name);
    }
};
