// Reader of the reference's .sim scene scripts: the parsing rules of
// SCRIPT_READER / SCRIPT_BLOCK (SCRIPT_READER.h, SCRIPT_READER.cpp), restated
// in standard C++ (std::regex in place of boost::regex).
//
//   - comments are blanked before parsing, in three passes
//     (SCRIPT_READER.cpp:298-381): "// NAME {...}" removes the whole braced
//     block; "//" removes to the end of the line, where the whitespace after
//     the slashes may run over line ends (so a bare "//" at a line end also
//     removes the next line); "/*" removes up to the first '/' that follows
//     at least one character (the reference's test of the closing "*/"
//     assigns instead of compares, :367);
//   - "#INCLUDE <file>" inserts a file once, resolved against the script's
//     directory (:232-296);
//   - top-level "NAME { ... }" blocks, nested blocks inside them; text
//     outside blocks is ignored (:383-421, :119-153);
//   - values "key = ..." of a block, its child blocks excluded, collected by
//     type in the order number, vector2, vector3, vector4, boolean, string;
//     a lookup takes the first value of that name in that order (:155-214,
//     :461-480);
//   - typed getters: integers by atoi, floats by atof, booleans true only for
//     the text "true", vectors parsed from "(a, b, c)" (0 for each component
//     when the value is not a vector), defaults when the key is absent
//     (:842-987);
//   - block paths "A.B.C" search child blocks by name, first match; a missing
//     block is an empty one (:482-507, :774-801).
#pragma once

#include <string>
#include <vector>

namespace fs {

struct ScriptValue {
    enum Type { NUMBER = 0, VECTOR2, VECTOR3, VECTOR4, BOOLEAN, STRING };
    std::string name;
    std::string val;
    Type type;
};

struct ScriptBlock {
    std::string name;
    std::vector<ScriptValue> values;
    std::vector<ScriptBlock> children;

    const ScriptValue* Value(const char* key) const;
    // child block by path "A.B" (relative to this block); an empty block if absent
    const ScriptBlock& Find(const char* path) const;

    int GetInteger(const char* key, int def = 0) const;
    double GetFloat(const char* key, double def = 0) const;
    bool GetBoolean(const char* key, bool def = false) const;
    std::string GetString(const char* key, const char* def = "Null") const;
    void GetVector3(const char* key, const double def[3], double out[3]) const;
    void GetInt3(const char* key, const int def[3], int out[3]) const;
};

class Script {
   public:
    // false (with a message) when the file cannot be read
    bool Load(const std::string& path, std::string* err);
    // top-level block by path "A.B.C"; an empty block if absent
    const ScriptBlock& Find(const char* path) const;
    const std::vector<ScriptBlock>& Blocks() const { return blocks_; }
    const std::vector<std::string>& Files() const { return files_; }
    const std::string& Directory() const { return dir_; }

    // the preprocessing and parsing steps, exposed for tests of the rules
    static void StripComments(std::string& text);
    static std::vector<std::string> CutBlocks(std::string& text);
    static ScriptBlock ParseBlock(const std::string& block_text);

   private:
    void ExpandIncludes(std::string& text);
    std::vector<ScriptBlock> blocks_;
    std::vector<std::string> files_;
    std::string dir_;
};

}  // namespace fs
