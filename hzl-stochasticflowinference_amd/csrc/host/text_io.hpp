// text_io.hpp -- the text encodings shared by the drop-in's file formats and
// the command-line tool: one number formatter (shortest-round-trip is NOT what
// the reference files use; they hold 17 significant digits, printf "%.17g"),
// a delimited-line writer, a token scanner with typed reads, and a small JSON
// object builder.  Internal to libhzl_gpu.so and hzl_amd_cli.
#pragma once

#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <system_error>
#include <type_traits>
#include <utility>
#include <vector>

namespace hzl::text {

// "%.17g": the digits std::ostream writes at setprecision(17).
inline std::string number(double v) {
    char buf[40];
    const auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::general, 17);
    return std::string(buf, r.ptr);
}
template <class I, std::enable_if_t<std::is_integral_v<I>, int> = 0>
inline std::string number(I v) {
    char buf[24];
    const auto r = std::to_chars(buf, buf + sizeof buf, v);
    return std::string(buf, r.ptr);
}
inline std::string number(bool v) {
    return v ? "1" : "0";
}
// std::to_string's spelling ("%f" for doubles), used where a value is shown
// the way a default option value or a summary statistic is shown.
inline std::string fixed6(double v) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "%f", v);
    return buf;
}

// A file of delimited lines.  Fields are appended with add(); end() closes
// the line.  A file that cannot be opened throws `open_failure` (by default
// `failure`), any later stream failure `failure`.
class LineWriter {
  public:
    LineWriter(const std::filesystem::path &path, char delimiter, std::string failure,
               const std::string &open_failure = {})
        : out_(path), delimiter_(delimiter), failure_(std::move(failure)) {
        if (!out_)
            throw std::runtime_error(open_failure.empty() ? failure_ : open_failure);
    }
    template <class T> LineWriter &add(const T &v) {
        if (open_)
            line_ += delimiter_;
        open_ = true;
        if constexpr (std::is_convertible_v<T, std::string_view>)
            line_ += std::string_view(v);
        else
            line_ += number(v);
        return *this;
    }
    LineWriter &skip() { // an empty field
        return add("");
    }
    void end() {
        line_ += '\n';
        out_.write(line_.data(), (std::streamsize)line_.size());
        line_.clear();
        open_ = false;
        if (!out_)
            throw std::runtime_error(failure_);
    }
    template <class... T> void line(const T &...v) {
        (add(v), ...);
        end();
    }
    void close() {
        out_.close();
        if (!out_)
            throw std::runtime_error(failure_);
    }

  private:
    std::ofstream out_;
    char delimiter_;
    std::string failure_, line_;
    bool open_ = false;
};

// Tokens of a text held in memory, separated by blanks and (optionally) commas.
class Scanner {
  public:
    Scanner(std::string text, bool commas) : text_(std::move(text)), commas_(commas) {}
    static std::string slurp(const std::filesystem::path &path) {
        std::ifstream in(path, std::ios::binary);
        return std::string(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    }
    // Next token (empty at the end of the text).
    std::string_view token() {
        while (pos_ < text_.size() && separator(text_[pos_]))
            ++pos_;
        const size_t begin = pos_;
        while (pos_ < text_.size() && !separator(text_[pos_]))
            ++pos_;
        return std::string_view(text_).substr(begin, pos_ - begin);
    }
    bool at_end() {
        while (pos_ < text_.size() && separator(text_[pos_]))
            ++pos_;
        return pos_ >= text_.size();
    }
    // Typed reads: false when the token is missing or does not parse whole.
    template <class T> bool read(T &out) {
        return parse(token(), out);
    }
    template <class T> static bool parse(std::string_view t, T &out) {
        if (t.empty())
            return false;
        if constexpr (std::is_same_v<T, bool>) {
            if (t != "0" && t != "1")
                return false;
            out = t == "1";
            return true;
        } else if constexpr (std::is_same_v<T, std::string>) {
            out.assign(t);
            return true;
        } else {
            // the numbers a stream extraction accepts: an optional sign, then a
            // digit (or a decimal point for floating point); no inf / nan words
            const char *b = t.data(), *e = t.data() + t.size();
            if (*b == '+')
                ++b;
            const char *d = b < e && *b == '-' ? b + 1 : b;
            const bool digit = d < e && ((*d >= '0' && *d <= '9') ||
                                         (std::is_floating_point_v<T> && *d == '.'));
            if (!digit || (std::is_unsigned_v<T> && d != b))
                return false;
            const auto r = std::from_chars(b, e, out);
            return r.ec == std::errc{} && r.ptr == e;
        }
    }

  private:
    bool separator(char c) const {
        return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f' ||
               (commas_ && c == ',');
    }
    std::string text_;
    bool commas_;
    size_t pos_ = 0;
};

// Split a text into lines (the terminating newline of each is dropped).
inline std::vector<std::string_view> lines_of(std::string_view text) {
    std::vector<std::string_view> out;
    size_t begin = 0;
    while (begin < text.size()) {
        size_t end = text.find('\n', begin);
        if (end == std::string_view::npos)
            end = text.size();
        out.push_back(text.substr(begin, end - begin));
        begin = end + 1;
    }
    return out;
}

// JSON string literal with the escapes the reference CLI writes.
inline std::string json_string(std::string_view v) {
    std::string s = "\"";
    for (unsigned char c : v) {
        if (c == '"' || c == '\\') {
            s += '\\';
            s += (char)c;
        } else if (c < 32) {
            char buf[8];
            std::snprintf(buf, sizeof buf, "\\u%04x", (unsigned)c);
            s += buf;
        } else {
            s += (char)c;
        }
    }
    return s + "\"";
}

// An ordered JSON object of scalar members, rendered on one line or one
// member per line with two-space indentation.
class JsonObject {
  public:
    JsonObject &raw(std::string key, std::string literal) {
        members_.emplace_back(std::move(key), std::move(literal));
        return *this;
    }
    template <class T> JsonObject &value(std::string key, const T &v) {
        if constexpr (std::is_convertible_v<T, std::string_view>)
            return raw(std::move(key), json_string(v));
        else
            return raw(std::move(key), number(v));
    }
    std::string compact() const {
        std::string s = "{";
        for (size_t i = 0; i < members_.size(); ++i)
            s += (i ? "," : "") + json_string(members_[i].first) + ":" + members_[i].second;
        return s + "}\n";
    }
    std::string pretty() const {
        std::string s = "{\n";
        for (size_t i = 0; i < members_.size(); ++i)
            s += "  " + json_string(members_[i].first) + ": " + members_[i].second +
                 (i + 1 < members_.size() ? ",\n" : "\n");
        return s + "}\n";
    }

  private:
    std::vector<std::pair<std::string, std::string>> members_;
};

inline void write_file(const std::filesystem::path &path, const std::string &text,
                       const std::string &failure) {
    std::ofstream out(path, std::ios::binary);
    out.write(text.data(), (std::streamsize)text.size());
    out.close();
    if (!out)
        throw std::runtime_error(failure);
}

} // namespace hzl::text
