/*
 * Utils.hpp -- host utilities of the APPFIS API: command-line flags
 * (`-name=value`, reference Appfis/Utils.hpp:225-283), string helpers
 * (Utils.hpp:47-120) and directory helpers (Utils.hpp:150-206).  Written against
 * C++17 (no std::ptr_fun / std::not1).
 */
#ifndef APPFIS_UTILS_HPP
#define APPFIS_UTILS_HPP

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include <dirent.h>
#include <strings.h>
#include <sys/stat.h>

namespace APPFIS
{
    typedef std::vector<std::string> StringVector;

    inline StringVector& Split(const std::string& s, char delim, StringVector& elems)
    {
        std::stringstream ss(s);
        std::string item;
        while (std::getline(ss, item, delim)) elems.push_back(item);
        return elems;
    }

    inline StringVector Split(const std::string& s, char delim)
    {
        StringVector elems;
        Split(s, delim, elems);
        return elems;
    }

    // converts every string to T with strtod (the reference goes through a stringstream)
    template <typename T>
    std::vector<T> VecstrToVectype(StringVector vs)
    {
        std::vector<T> out;
        out.reserve(vs.size());
        for (const std::string& s : vs) {
            std::stringstream ss(s);
            T v{};
            ss >> v;
            out.push_back(v);
        }
        return out;
    }

    static inline std::string& LeftTrim(std::string& s)
    {
        s.erase(s.begin(), std::find_if(s.begin(), s.end(), [](unsigned char c) { return !std::isspace(c); }));
        return s;
    }

    static inline std::string& RightTrim(std::string& s)
    {
        s.erase(std::find_if(s.rbegin(), s.rend(), [](unsigned char c) { return !std::isspace(c); }).base(), s.end());
        return s;
    }

    static inline std::string& Trim(std::string& s) { return LeftTrim(RightTrim(s)); }

    inline std::string GetParentDir(const char* path)
    {
        std::string p(path);
        size_t found = p.find_last_of("/\\");
        if (found != std::string::npos && found == p.length() - 1) p = p.substr(0, found);
        found = p.find_last_of("/\\");
        return found == std::string::npos ? std::string() : p.substr(0, found);
    }

    // creates `path` when it names a directory (no extension after the last slash)
    inline void CreateDir(std::string path)
    {
        if (path.empty()) return;
        size_t dot = path.find_last_of('.'), slash = path.find_last_of("/\\");
        if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) return;
        DIR* dir = opendir(path.c_str());
        if (dir) {
            closedir(dir);
            return;
        }
        mkdir(path.c_str(), S_IRWXU);
    }

    inline int StringRemoveDelimiter(char delimiter, const char* string)
    {
        int start = 0;
        while (string[start] == delimiter) start++;
        if (start >= static_cast<int>(strlen(string)) - 1) return 0;
        return start;
    }

    // true when argv holds `-name` or `-name=...` (case-insensitive)
    inline bool CheckArgument(int argc, char* argv[], const char* name)
    {
        const int len = static_cast<int>(strlen(name));
        for (int i = 1; i < argc; i++) {
            const char* a = &argv[i][StringRemoveDelimiter('-', argv[i])];
            const char* eq = strchr(a, '=');
            const int alen = static_cast<int>(eq ? eq - a : strlen(a));
            if (alen == len && strncasecmp(a, name, len) == 0) return true;
        }
        return false;
    }

    // integer value of `-name=value` (0 when absent or bare)
    inline int GetArgument(int argc, char* argv[], const char* name)
    {
        const int len = static_cast<int>(strlen(name));
        bool found = false;
        int value = -1;
        for (int i = 1; i < argc; i++) {
            const char* a = &argv[i][StringRemoveDelimiter('-', argv[i])];
            if (strncasecmp(a, name, len) == 0) {
                if (len + 1 <= static_cast<int>(strlen(a))) value = atoi(&a[len + (a[len] == '=' ? 1 : 0)]);
                else value = 0;
                found = true;
            }
        }
        return found ? value : 0;
    }
}  // namespace APPFIS

#endif
