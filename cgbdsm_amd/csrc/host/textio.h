// Minimal fast tokenizer for the `.graph` / update-stream text formats
// (README.md:151-192 of the reference): whitespace separated tokens, the
// first token of a record is a one-character type tag.
#pragma once
#include <cstdio>
#include <string>
#include <vector>

#include "common.h"

namespace cg {

class TextReader {
public:
    explicit TextReader(const std::string& path) {
        FILE* f = fopen(path.c_str(), "rb");
        if (!f) throw Error("Failed to open: " + path);
        fseek(f, 0, SEEK_END);
        long sz = ftell(f);
        fseek(f, 0, SEEK_SET);
        buf_.resize(sz > 0 ? (size_t)sz : 0);
        if (sz > 0 && fread(buf_.data(), 1, (size_t)sz, f) != (size_t)sz) {
            fclose(f);
            throw Error("short read: " + path);
        }
        fclose(f);
    }
    bool NextChar(char* c) {
        SkipWs();
        if (p_ >= buf_.size()) return false;
        *c = buf_[p_++];
        return true;
    }
    long Int() {
        SkipWs();
        if (p_ >= buf_.size()) return 0;
        bool neg = false;
        if (buf_[p_] == '-' || buf_[p_] == '+') neg = buf_[p_++] == '-';
        long v = 0;
        while (p_ < buf_.size() && buf_[p_] >= '0' && buf_[p_] <= '9') v = v * 10 + (buf_[p_++] - '0');
        return neg ? -v : v;
    }
    void SkipLine() {
        while (p_ < buf_.size() && buf_[p_] != '\n') p_++;
    }

private:
    void SkipWs() {
        while (p_ < buf_.size() && (buf_[p_] == ' ' || buf_[p_] == '\n' || buf_[p_] == '\r' || buf_[p_] == '\t')) p_++;
    }
    std::vector<char> buf_;
    size_t p_ = 0;
};

}  // namespace cg
