// Case-dictionary reader with the reference's lookup rules (src/dictIO):
//  * a file is read as its whitespace-separated words joined by single spaces
//    (openDict.c:33-71), so comments stay in the text;
//  * "a/b/c" descends into braced blocks a then b: the FIRST occurrence of the name
//    anywhere in the text that is followed (after spaces) by '{' wins
//    (bracedEntry.c:13-141, lookUpEntry.c:32-79);
//  * the last component is a plain substring match: the value is the text from one
//    character after the first occurrence of the name up to the next ';'
//    (singleEntry.c:14-64), so a key also matches inside a longer word;
//  * scalars go through strtod (lookUpScalarEntry.c:30), integers through atoi,
//    vectors are "( ... )" lists (lookUpVectorEntry.c:28-94).
#pragma once

#include <string>
#include <vector>

namespace mplb {

class Dict {
public:
    bool open(const std::string& path);
    bool lookup(const std::string& ename, std::string* out) const;
    bool scalar(const std::string& ename, double df, double* out) const;
    bool integer(const std::string& ename, int df, int* out) const;
    bool vector(const std::string& ename, std::vector<double>* out) const;
    bool string(const std::string& ename, const std::string& df, std::string* out) const;
    // lookUpStringList.c: every component of ename is a braced block; the words of the
    // innermost block are the list
    bool string_list(const std::string& ename, std::vector<std::string>* out) const;
    const std::string& text() const { return content_; }
    void set_text(const std::string& t) { content_ = t; }

private:
    std::string content_;
};

}  // namespace mplb
