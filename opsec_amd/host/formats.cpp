// Config and .abn formats (see formats.h).
#include "formats.h"

#include <getopt.h>

#include <cctype>
#include <cstdarg>
#include <cstring>

namespace opsec_host {

bool KeyLess::operator()(const std::string& a, const std::string& b) const {
    return std::strcmp(a.c_str(), b.c_str()) < 0;
}

static std::string trim(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) b++;
    while (e > b && std::isspace((unsigned char)s[e - 1])) e--;
    return s.substr(b, e - b);
}

bool Config::read_line(const std::string& line) {
    std::string t = trim(line);
    if (t.empty() || t[0] == '#') return true;
    size_t eq = t.find('=');
    if (eq == std::string::npos || eq == 0) {
        std::fprintf(stderr, "cfg_read_line: '%s' is not a valid configuration line\n", line.c_str());
        return false;
    }
    std::string key = trim(t.substr(0, eq));
    std::string val = trim(t.substr(eq + 1));
    if (val.size() >= 2 && ((val.front() == '"' && val.back() == '"') || (val.front() == '\'' && val.back() == '\'')))
        val = val.substr(1, val.size() - 2);
    kv[key] = val;
    return true;
}

bool Config::read_file(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) {
        std::fprintf(stderr, "cfg_read_file: could not read file '%s'\n", path.c_str());
        return false;
    }
    char buf[4096];
    bool ok = true;
    while (std::fgets(buf, sizeof buf, f))
        if (!read_line(buf)) {
            ok = false;
            break;
        }
    std::fclose(f);
    return ok;
}

static std::vector<std::string> split_keys(const std::string& s) {
    std::vector<std::string> out;
    size_t p = 0;
    while (p <= s.size()) {
        size_t c = s.find(',', p);
        if (c == std::string::npos) c = s.size();
        std::string k = trim(s.substr(p, c - p));
        if (!k.empty()) out.push_back(k);
        p = c + 1;
    }
    return out;
}

bool Config::has_all(const std::string& keys) const { return missing(keys).empty(); }

std::string Config::missing(const std::string& keys) const {
    std::string m;
    for (const auto& k : split_keys(keys))
        if (!has(k)) m += "," + k;
    return m;
}

std::string Config::get(const std::string& key) const {
    auto it = kv.find(key);
    return it == kv.end() ? std::string() : it->second;
}

bool Config::get_int(const std::string& key, int* v) const {
    return std::sscanf(get(key).c_str(), "%i", v) == 1;
}

bool Config::get_double(const std::string& key, double* v) const {
    return std::sscanf(get(key).c_str(), "%lf", v) == 1;
}

std::vector<double> Config::get_doubles(const std::string& key, int nmax) const {
    std::vector<double> out;
    std::string s = get(key);
    size_t p = 0;
    while ((int)out.size() < nmax && p < s.size()) {
        while (p < s.size() && (s[p] == ' ' || s[p] == ',')) p++;
        if (p >= s.size()) break;
        size_t e = p;
        while (e < s.size() && s[e] != ' ' && s[e] != ',') e++;
        double v;
        if (std::sscanf(s.substr(p, e - p).c_str(), "%lf", &v) == 1) out.push_back(v);
        p = e;
    }
    return out;
}

int Config::get_enum(const std::string& key, const std::vector<std::pair<std::string, int>>& names, int def) const {
    std::string s = get(key);
    for (const auto& nv : names)
        if (nv.first.compare(0, s.size(), s) == 0 && s.size() <= nv.first.size()) return nv.second;
    return def;
}

void Config::set_int(const std::string& key, int v) { kv[key] = std::to_string(v); }

void Config::set_double(const std::string& key, double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g", v);
    kv[key] = buf;
}

Config Config::sub(const std::string& prefix) const {
    Config c;
    for (const auto& e : kv)
        if (e.first.compare(0, prefix.size(), prefix) == 0) c.kv[e.first.substr(prefix.size())] = e.second;
    return c;
}

void Config::write(FILE* f) const {
    for (const auto& e : kv) {
        const std::string& v = e.second;
        bool q = !v.empty() && (std::isspace((unsigned char)v.front()) || std::isspace((unsigned char)v.back()));
        if (q) std::fprintf(f, "%s = \"%s\"\n", e.first.c_str(), v.c_str());
        else std::fprintf(f, "%s = %s\n", e.first.c_str(), v.c_str());
    }
}

bool parse_command_line(int argc, char** argv, const char* usage, Config& cfg, int* exit_code) {
    optind = 1;
    int opt;
    while ((opt = getopt(argc, argv, "hc:")) != -1) {
        switch (opt) {
            case 'h':
                std::fputs(usage, stdout);
                *exit_code = 0;
                return false;
            case 'c':
                cfg.read_file(optarg);
                break;
            default:
                std::fputs(usage, stdout);
                *exit_code = 1;
                return false;
        }
    }
    for (int i = optind; i < argc; i++) cfg.read_line(argv[i]);
    return true;
}

// ---- .abn -----------------------------------------------------------------

int abn_calcsize(const std::string& fmt) {
    int size = 0, num = 0;
    bool have_num = false;
    for (char c : fmt) {
        if (std::isdigit((unsigned char)c)) {
            num = num * 10 + (c - '0');
            have_num = true;
            continue;
        }
        int len;
        switch (c) {
            case 'x': case 'c': case 'b': case 'B': case '?': case 's': case 'p': len = 1; break;
            case 'h': case 'H': len = 2; break;
            case 'i': case 'I': case 'f': len = 4; break;
            case 'l': case 'L': len = (int)sizeof(long); break;
            case 'q': case 'Q': case 'd': len = 8; break;
            case ' ': continue;
            case '@': case '=': case '<': case '>': case '!': continue;
            default: return -1;
        }
        size += (have_num ? num : 1) * len;
        num = 0;
        have_num = false;
    }
    return size;
}

char abn_endian() {
    unsigned int x = 1;
    return *(unsigned char*)&x ? 'L' : 'B';
}

void abn_write_header(FILE* f, size_t n, const std::string& fmt, const Config* opts) {
    std::string clean;
    for (char c : fmt)
        if (!std::isspace((unsigned char)c)) clean += c;
    std::fprintf(f, "\n");
    if (opts) opts->write(f);
    std::fprintf(f, "#!Begin data block: n = %zu, size = %i, endian = %c, format = %s\n", n, abn_calcsize(clean),
                 abn_endian(), clean.c_str());
}

size_t abn_write(FILE* f, const void* data, size_t n, const std::string& fmt, const Config* opts) {
    abn_write_header(f, n, fmt, opts);
    return std::fwrite(data, (size_t)abn_calcsize(fmt), n, f);
}

bool abn_read_header(FILE* f, AbnHeader* h) {
    char line[1024];
    while (std::fgets(line, sizeof line, f)) {
        size_t n, size;
        char endian;
        char fmt[128];
        if (std::sscanf(line, "#!Begin data block: n = %zu, size = %zu, endian = %c, format = %127s", &n, &size,
                        &endian, fmt) == 4) {
            h->n = n;
            h->size = size;
            h->endian = endian;
            h->fmt = fmt;
            return true;
        }
        h->opts.read_line(line);
    }
    std::fprintf(stderr, "abn_read_header: no header found\n");
    return false;
}

void abn_byte_swap(char* data, size_t n, const std::string& fmt) {
    std::vector<int> fields;
    int num = 0;
    bool have = false;
    for (char c : fmt) {
        if (std::isdigit((unsigned char)c)) {
            num = num * 10 + (c - '0');
            have = true;
            continue;
        }
        int len = abn_calcsize(std::string(1, c));
        if (len <= 0) continue;
        for (int i = 0; i < (have ? num : 1); i++) fields.push_back(len);
        num = 0;
        have = false;
    }
    for (size_t i = 0; i < n; i++)
        for (int len : fields) {
            for (int a = 0, b = len - 1; a < b; a++, b--) std::swap(data[a], data[b]);
            data += len;
        }
}

void info(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stdout, fmt, ap);
    va_end(ap);
    std::fflush(stdout);
}

void error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stderr, fmt, ap);
    va_end(ap);
    std::fflush(stderr);
}

bool SplitReader::open(const std::string& path) {
    close();
    cur_ = path;
    index_ = 0;
    f_ = std::fopen(path.c_str(), "rb");
    return f_ != nullptr;
}

void SplitReader::close() {
    if (f_) std::fclose(f_);
    f_ = nullptr;
}

void SplitReader::open_next() {
    if (!f_) return;
    std::fclose(f_);
    f_ = nullptr;
    index_++;
    if (cur_.size() < 2 || index_ > 99) return;
    cur_[cur_.size() - 2] = (char)('0' + index_ / 10);
    cur_[cur_.size() - 1] = (char)('0' + index_ % 10);
    f_ = std::fopen(cur_.c_str(), "rb");
}

size_t SplitReader::read(void* buf, size_t bytes) {
    char* p = (char*)buf;
    size_t done = 0;
    while (done < bytes && f_ && !std::ferror(f_)) {
        size_t n = std::fread(p + done, 1, bytes - done, f_);
        done += n;
        if (n == 0 && std::feof(f_)) open_next();
    }
    return done;
}

}  // namespace opsec_host
