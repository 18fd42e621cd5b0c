// Minimal reader for the LucasNTT.ini key surface and the stats file that doubles as a
// resume file.  Behaviour follows what the reference relies on from SimpleIni
// (main_process.cu:69-123): keys outside any [section] live in the unnamed section,
// key lookup ignores case, surrounding blanks are trimmed, '#' and ';' start comment
// lines, a later duplicate key replaces an earlier one, integers accept a 0x prefix.
#pragma once
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>

class Ini {
   public:
    bool LoadFile(const std::string &path) {
        std::ifstream f(path);
        if (!f) return false;
        std::string line, section;
        while (std::getline(f, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            std::string t = trim(line);
            if (t.empty() || t[0] == '#' || t[0] == ';') continue;
            if (t[0] == '[') {
                size_t e = t.find(']');
                section = lower(trim(t.substr(1, e == std::string::npos ? std::string::npos : e - 1)));
                continue;
            }
            size_t eq = t.find('=');
            if (eq == std::string::npos) continue;
            std::string k = lower(trim(t.substr(0, eq)));
            std::string v = trim(t.substr(eq + 1));
            values_[section + "\x1f" + k] = v;
        }
        loaded_ = true;
        return true;
    }
    bool Loaded() const { return loaded_; }
    std::string GetValue(const std::string &key, const std::string &def) const {
        auto it = values_.find(std::string("\x1f") + lower(key));
        return it == values_.end() ? def : it->second;
    }
    long GetLongValue(const std::string &key, long def) const {
        auto it = values_.find(std::string("\x1f") + lower(key));
        if (it == values_.end() || it->second.empty()) return def;
        const char *s = it->second.c_str();
        char *end = nullptr;
        long v = std::strtol(s, &end, 0);
        return end == s ? def : v;
    }

   private:
    static std::string trim(const std::string &s) {
        size_t a = s.find_first_not_of(" \t"), b = s.find_last_not_of(" \t");
        return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    }
    static std::string lower(std::string s) {
        for (auto &c : s) c = (char)std::tolower((unsigned char)c);
        return s;
    }
    std::map<std::string, std::string> values_;
    bool loaded_ = false;
};
