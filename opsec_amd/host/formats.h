// Config files and annotated binary (.abn) files, with the reference's
// semantics (src/cfg.c, src/abn.c, src/opsec.c), for the klt / comma /
// estimate drop-in programs.
#pragma once

#include <cstdio>
#include <map>
#include <string>
#include <vector>

namespace opsec_host {

// Byte-wise ordering of keys, as strcmp sorts them (src/cfg.c:49-120).
struct KeyLess {
    bool operator()(const std::string& a, const std::string& b) const;
};

class Config {
public:
    // Parse one "key = value" line: leading/trailing whitespace trimmed, '#'
    // comments and blank lines ignored, matching quotes stripped from values.
    // Returns false (and reports) on a line without '=' or with an empty key.
    bool read_line(const std::string& line);
    bool read_file(const std::string& path);
    bool has(const std::string& key) const { return kv.count(key) != 0; }
    bool has_all(const std::string& comma_keys) const;
    std::string missing(const std::string& comma_keys) const;   // ",a,b" of missing keys, "" if none
    std::string get(const std::string& key) const;               // "" if missing
    bool get_int(const std::string& key, int* v) const;          // sscanf %i semantics
    bool get_double(const std::string& key, double* v) const;
    std::vector<double> get_doubles(const std::string& key, int nmax) const;
    // value must be a prefix of one of the names (src/cfg.c:486-501); `def` if none
    int get_enum(const std::string& key, const std::vector<std::pair<std::string, int>>& names, int def) const;
    void set(const std::string& key, const std::string& value) { kv[key] = value; }
    void set_int(const std::string& key, int v);
    void set_double(const std::string& key, double v);            // %g
    Config sub(const std::string& prefix) const;
    void write(FILE* f) const;                                    // alphabetical "key = value"
    std::map<std::string, std::string, KeyLess> kv;
};

// Command line of the reference programs: prog [-h] [-c FILE] key=value ...
// (src/opsec.c:19-56).  Returns false and sets exit_code when the program
// must exit (help, bad switch).
bool parse_command_line(int argc, char** argv, const char* usage, Config& cfg, int* exit_code);

// ---- .abn ----------------------------------------------------------------
struct AbnHeader {
    size_t n = 0, size = 0;
    char endian = 'L';
    std::string fmt;
    Config opts;
};

int abn_calcsize(const std::string& fmt);   // struct-module size, -1 if invalid
char abn_endian();
// "\n", opts (if any), the data tag line (src/abn.c:186-211)
void abn_write_header(FILE* f, size_t n, const std::string& fmt, const Config* opts);
size_t abn_write(FILE* f, const void* data, size_t n, const std::string& fmt, const Config* opts);
// scan to the tag line, feeding other lines to opts (src/abn.c:221-247)
bool abn_read_header(FILE* f, AbnHeader* h);
void abn_byte_swap(char* data, size_t n, const std::string& fmt);

// ---- files cut into pieces ------------------------------------------------
// Sequential reader over a file that may have been cut into pieces by
// `split -d -a 2`: when one piece ends, the next is the current name with its
// last two characters replaced by the piece number 01, 02, ... (SplitFile,
// src/SplitFile.cpp:77-170).  The abn header is read from stream(); data
// reads may then run on into later pieces.  A file that was never split reads
// exactly as with fread.
class SplitReader {
public:
    SplitReader() = default;
    SplitReader(const SplitReader&) = delete;
    SplitReader& operator=(const SplitReader&) = delete;
    ~SplitReader() { close(); }
    bool open(const std::string& path);
    void close();
    FILE* stream() const { return f_; }
    const std::string& filename() const { return cur_; }
    size_t read(void* buf, size_t bytes);   // bytes actually read
private:
    void open_next();
    std::string cur_;
    int index_ = 0;
    FILE* f_ = nullptr;
};

// ---- logging --------------------------------------------------------------
void info(const char* fmt, ...);
void error(const char* fmt, ...);

}  // namespace opsec_host
