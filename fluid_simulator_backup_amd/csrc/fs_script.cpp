// Reader of .sim scene scripts; the rules are listed in fs_script.h.
#include "fs_script.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <regex>

namespace fs {

namespace {

// value patterns of SCRIPT_READER.cpp:7-17
const std::regex& value_regex(ScriptValue::Type t) {
    static const std::regex number(R"RE(\s*(\w+)\s*=\s*(-*\d*\.*\d+)\s*)RE");
    static const std::regex vector2(R"RE(\s*(\w+)\s*=\s*(\(\s*-*\d*\.*\d+\s*,\s*-*\d*\.*\d+\s*\))\s*)RE");
    static const std::regex vector3(
        R"RE(\s*(\w+)\s*=\s*(\(\s*-*\d*\.*\d+\s*,\s*-*\d*\.*\d+\s*,\s*-*\d*\.*\d+\s*\))\s*)RE");
    static const std::regex vector4(
        R"RE(\s*(\w+)\s*=\s*(\(\s*-*\d*\.*\d+\s*,\s*-*\d*\.*\d+\s*,\s*-*\d*\.*\d+\s*,\s*-*\d*\.*\d+\s*\))\s*)RE");
    static const std::regex boolean(R"RE(\s*(\w+)\s*=\s*(true|false)\s*)RE");
    static const std::regex string(R"RE(\s*(\w+)\s*=\s*"(\s*[:\\.\w/\s\-*]+\s*)"\s*)RE");
    switch (t) {
        case ScriptValue::NUMBER: return number;
        case ScriptValue::VECTOR2: return vector2;
        case ScriptValue::VECTOR3: return vector3;
        case ScriptValue::VECTOR4: return vector4;
        case ScriptValue::BOOLEAN: return boolean;
        default: return string;
    }
}

void blank(std::string& s, size_t b, size_t e) {
    for (size_t i = b; i < e && i < s.size(); i++) s[i] = ' ';
}

// first match of re in s at or after pos: [b, e) and the submatches
bool search_from(const std::string& s, size_t pos, const std::regex& re, std::smatch& m, size_t& b, size_t& e) {
    if (pos > s.size()) return false;
    if (!std::regex_search(s.cbegin() + (long)pos, s.cend(), m, re)) return false;
    b = pos + (size_t)m.position(0);
    e = b + (size_t)m.length(0);
    return true;
}

// path "A.B.C" split at dots; a dot in the first character does not split
// (TrimBlockNamePath, SCRIPT_READER.cpp:423-438)
std::vector<std::string> split_path(const char* path) {
    std::vector<std::string> out;
    std::string p(path);
    size_t begin = 0;
    for (size_t i = 1; i < p.size(); i++) {
        if (p[i] == '.') {
            out.push_back(p.substr(begin, i - begin));
            begin = i + 1;
        }
    }
    out.push_back(p.size() > begin ? p.substr(begin) : std::string());
    return out;
}

const ScriptBlock& empty_block() {
    static const ScriptBlock e;
    return e;
}

const ScriptBlock* search(const std::vector<ScriptBlock>& v, const std::string& name) {
    for (const ScriptBlock& b : v)
        if (b.name == name) return &b;
    return nullptr;
}

const ScriptBlock& find_path(const std::vector<ScriptBlock>& top, const char* path) {
    std::vector<std::string> parts = split_path(path);
    const ScriptBlock* cur = search(top, parts[0]);
    for (size_t i = 1; cur && i < parts.size(); i++) cur = search(cur->children, parts[i]);
    return cur ? *cur : empty_block();
}

}  // namespace

// ---- comments (SCRIPT_READER.cpp:298-381) --------------------------------
void Script::StripComments(std::string& t) {
    std::smatch m;
    size_t b, e, pos;
    // "// NAME {" up to its matching brace
    static const std::regex block_comment(R"RE(\s*//\s*\w*\s*\{\s*)RE");
    pos = 0;
    while (search_from(t, pos, block_comment, m, b, e)) {
        int depth = 1;
        while (depth != 0 && e < t.size()) {
            if (t[e] == '}') depth--;
            if (t[e] == '{') depth++;
            e++;
        }
        blank(t, b, e);
        pos = e;
    }
    // "//" to the end of the line; the scan starts one character past the
    // whitespace that follows the slashes
    static const std::regex line_comment(R"RE(\s*//\s*)RE");
    pos = 0;
    while (search_from(t, pos, line_comment, m, b, e)) {
        while (e < t.size()) {
            e++;
            if (e >= t.size() || t[e] == '\n') break;
        }
        blank(t, b, e);
        pos = e;
    }
    // "/*" up to and including the first '/' preceded by a character of the
    // comment
    static const std::regex c_comment(R"RE(\s*/\*\s*)RE");
    pos = 0;
    while (search_from(t, pos, c_comment, m, b, e)) {
        while (e < t.size()) {
            if (e + 1 < t.size() && t[e + 1] == '/') {
                e += 2;
                break;
            }
            e++;
        }
        blank(t, b, e);
        pos = e;
    }
}

// ---- #INCLUDE (SCRIPT_READER.cpp:232-296) ---------------------------------
void Script::ExpandIncludes(std::string& t) {
    static const std::regex include(R"RE(\s*#INCLUDE\s*<\s*([\.\w\/]+)\s*>\s*)RE");
    std::smatch m;
    size_t b, e, pos = 0;
    while (search_from(t, pos, include, m, b, e)) {
        std::string name = m[1].str();
        bool seen = false;
        for (const std::string& f : files_) seen = seen || f == name;
        if (seen) {
            blank(t, b, e);
            pos = b;
            continue;
        }
        std::string full = (!name.empty() && name[0] == '/') || dir_.empty() ? name : dir_ + "/" + name;
        FILE* f = std::fopen(full.c_str(), "rb");
        if (!f) {
            blank(t, b, e);
            pos = e;
            continue;
        }
        files_.push_back(name);
        std::string inc;
        int c;
        while ((c = std::fgetc(f)) != EOF) inc += (char)c;
        std::fclose(f);
        t = t.substr(0, b) + '\n' + inc + '\n' + t.substr(e);
        StripComments(t);
        pos = b;
    }
}

// ---- blocks (SCRIPT_READER.cpp:383-421, :119-153) --------------------------
std::vector<std::string> Script::CutBlocks(std::string& t) {
    static const std::regex head(R"RE((\w+)\s*\{\s*)RE");
    std::vector<std::string> out;
    std::smatch m;
    size_t b, e, pos = 0;
    while (search_from(t, pos, head, m, b, e)) {
        int depth = 1;
        while (depth != 0) {
            if (e >= t.size()) break;
            if (t[e] == '{') depth++;
            else if (t[e] == '}') depth--;
            e++;
        }
        out.push_back(t.substr(b, e - b));
        blank(t, b, e);
        pos = e;
    }
    return out;
}

ScriptBlock Script::ParseBlock(const std::string& text) {
    // "\s*(\w+)\s*\{\s*(.*)\s*\}\s*": the name, and everything from the first
    // brace (and the whitespace after it) to the last closing brace
    static const std::regex head(R"RE(\s*(\w+)\s*\{\s*)RE");
    ScriptBlock blk;
    std::smatch m;
    size_t b, e;
    std::string data;
    size_t last = text.rfind('}');
    if (search_from(text, 0, head, m, b, e) && last != std::string::npos && last + 1 > e) {
        blk.name = m[1].str();
        data = text.substr(e, last - e);
    }
    std::vector<std::string> kids = CutBlocks(data);
    for (int t = ScriptValue::NUMBER; t <= ScriptValue::STRING; t++) {
        const std::regex& re = value_regex((ScriptValue::Type)t);
        size_t pos = 0;
        while (search_from(data, pos, re, m, b, e)) {
            blk.values.push_back(ScriptValue{m[1].str(), m[2].str(), (ScriptValue::Type)t});
            if (e == pos && e == b) e++;  // an empty match cannot occur (\w+), kept for safety
            pos = e;
        }
    }
    for (const std::string& k : kids) blk.children.push_back(ParseBlock(k));
    return blk;
}

bool Script::Load(const std::string& path, std::string* err) {
    blocks_.clear();
    files_.clear();
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        if (err) *err = "cannot open script " + path;
        return false;
    }
    size_t slash = path.find_last_of('/');
    dir_ = slash == std::string::npos ? std::string() : path.substr(0, slash);
    files_.push_back(path);
    std::string text;
    int c;
    while ((c = std::fgetc(f)) != EOF) text += (char)c;
    std::fclose(f);
    StripComments(text);
    ExpandIncludes(text);
    for (const std::string& b : CutBlocks(text)) blocks_.push_back(ParseBlock(b));
    return true;
}

const ScriptBlock& Script::Find(const char* path) const { return find_path(blocks_, path); }

// ---- block lookups and typed getters (SCRIPT_READER.cpp:643-662, 774-987) --
const ScriptValue* ScriptBlock::Value(const char* key) const {
    for (const ScriptValue& v : values)
        if (v.name == key) return &v;
    return nullptr;
}

const ScriptBlock& ScriptBlock::Find(const char* path) const { return find_path(children, path); }

int ScriptBlock::GetInteger(const char* key, int def) const {
    const ScriptValue* v = Value(key);
    return v ? std::atoi(v->val.c_str()) : def;
}

double ScriptBlock::GetFloat(const char* key, double def) const {
    const ScriptValue* v = Value(key);
    return v ? std::atof(v->val.c_str()) : def;
}

bool ScriptBlock::GetBoolean(const char* key, bool def) const {
    const ScriptValue* v = Value(key);
    return v ? v->val == "true" : def;
}

std::string ScriptBlock::GetString(const char* key, const char* def) const {
    const ScriptValue* v = Value(key);
    return v ? v->val : std::string(def);
}

namespace {
const std::regex& vector3_value() {
    static const std::regex re(
        R"RE(\s*\(\s*(-*\d*\.*\d+)\s*,\s*(-*\d*\.*\d+)\s*,\s*(-*\d*\.*\d+)\s*\)\s*)RE");
    return re;
}
}  // namespace

void ScriptBlock::GetVector3(const char* key, const double def[3], double out[3]) const {
    const ScriptValue* v = Value(key);
    if (!v) {
        for (int a = 0; a < 3; a++) out[a] = def[a];
        return;
    }
    std::smatch m;
    bool ok = std::regex_match(v->val, m, vector3_value());
    for (int a = 0; a < 3; a++) out[a] = ok ? std::atof(m[a + 1].str().c_str()) : 0.0;
}

void ScriptBlock::GetInt3(const char* key, const int def[3], int out[3]) const {
    const ScriptValue* v = Value(key);
    if (!v) {
        for (int a = 0; a < 3; a++) out[a] = def[a];
        return;
    }
    std::smatch m;
    bool ok = std::regex_match(v->val, m, vector3_value());
    for (int a = 0; a < 3; a++) out[a] = ok ? std::atoi(m[a + 1].str().c_str()) : 0;
}

}  // namespace fs
