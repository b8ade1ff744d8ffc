// Reader of OpenPV's .params language for the C++ host (see pvparams.h).
// The tokens and statements follow src/io/parser/params.l and params.y; the
// group / include / overwrite semantics follow PVParams's action_* callbacks
// (src/io/PVParams.hpp:381-399).
#include "pvparams.h"

#include <cctype>
#include <cfloat>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::string gError;

struct Value {
   enum Kind { NUMBER, STRING, NULLSTR, ARRAY } kind = NUMBER;
   double number = 0.0;
   std::string text;
   std::vector<double> array;
};

struct Param {
   std::string key;
   Value value;
};

struct Group {
   std::string keyword, name;
   std::vector<Param> params;
   Param *find(const std::string &key) {
      for (auto &p : params) {
         if (p.key == key) {
            return &p;
         }
      }
      return nullptr;
   }
};

enum Tok { T_END, T_ID, T_OVERWRITE, T_NUMBER, T_STRING, T_NULL, T_INCLUDE, T_SYM };

struct Token {
   Tok kind;
   std::string text;
   double number = 0.0;
   int line      = 0;
};

bool isIdStart(char c) { return std::isalpha((unsigned char)c) != 0; }
bool isIdChar(char c) { return std::isalnum((unsigned char)c) != 0 || c == '_'; }

// Tokenizer: numbers (strtod forms), identifiers, @identifiers, quoted
// strings, the keywords true/false/infinity/NULL/#include, single-character
// symbols; // comments and whitespace skipped.
bool tokenize(const std::string &s, std::vector<Token> &out) {
   size_t i = 0;
   int line = 1;
   while (i < s.size()) {
      char c = s[i];
      if (c == '\n') {
         line++;
         i++;
         continue;
      }
      if (std::isspace((unsigned char)c)) {
         i++;
         continue;
      }
      if (c == '/' && i + 1 < s.size() && s[i + 1] == '/') {
         while (i < s.size() && s[i] != '\n') {
            i++;
         }
         continue;
      }
      Token t;
      t.line = line;
      if (s.compare(i, 8, "#include") == 0) {
         t.kind = T_INCLUDE;
         i += 8;
         out.push_back(t);
         continue;
      }
      if (c == '"') {
         size_t j = s.find('"', i + 1);
         if (j == std::string::npos) {
            gError = "line " + std::to_string(line) + ": unterminated string";
            return false;
         }
         t.kind = T_STRING;
         t.text = s.substr(i + 1, j - i - 1);
         out.push_back(t);
         i = j + 1;
         continue;
      }
      if (c == '@' && i + 1 < s.size() && isIdStart(s[i + 1])) {
         size_t j = i + 1;
         while (j < s.size() && isIdChar(s[j])) {
            j++;
         }
         t.kind = T_OVERWRITE;
         t.text = s.substr(i + 1, j - i - 1);
         out.push_back(t);
         i = j;
         continue;
      }
      bool sign = (c == '+' || c == '-');
      size_t k  = sign ? i + 1 : i;
      if (k < s.size() && (std::isdigit((unsigned char)s[k]) || (s[k] == '.' && k + 1 < s.size()
                                                                      && std::isdigit((unsigned char)s[k + 1])))) {
         char *end = nullptr;
         double v  = std::strtod(s.c_str() + i, &end);
         t.kind    = T_NUMBER;
         t.number  = v;
         out.push_back(t);
         i = (size_t)(end - s.c_str());
         continue;
      }
      if (sign && s.compare(k, 8, "infinity") == 0) {
         t.kind   = T_NUMBER;
         t.number = c == '-' ? -FLT_MAX : FLT_MAX;
         out.push_back(t);
         i = k + 8;
         continue;
      }
      if (isIdStart(c)) {
         size_t j = i;
         while (j < s.size() && isIdChar(s[j])) {
            j++;
         }
         std::string id = s.substr(i, j - i);
         if (id == "true" || id == "false") {
            t.kind   = T_NUMBER;
            t.number = id == "true" ? 1.0 : 0.0;
         }
         else if (id == "infinity") {
            t.kind   = T_NUMBER;
            t.number = FLT_MAX;
         }
         else if (id == "NULL") {
            t.kind = T_NULL;
         }
         else {
            t.kind = T_ID;
            t.text = id;
         }
         out.push_back(t);
         i = j;
         continue;
      }
      if (std::strchr("{}[]=,;:", c)) {
         t.kind = T_SYM;
         t.text = std::string(1, c);
         out.push_back(t);
         i++;
         continue;
      }
      gError = "line " + std::to_string(line) + ": unexpected character '" + std::string(1, c) + "'";
      return false;
   }
   Token e;
   e.kind = T_END;
   e.line = line;
   out.push_back(e);
   return true;
}

struct Parser {
   std::vector<Token> toks;
   size_t pos = 0;
   std::vector<Group> groups;
   std::vector<Param> directives;

   const Token &peek() const { return toks[pos]; }
   bool fail(const std::string &msg) {
      gError = "line " + std::to_string(peek().line) + ": " + msg;
      return false;
   }
   bool sym(const char *c) {
      if (peek().kind == T_SYM && peek().text == c) {
         pos++;
         return true;
      }
      return false;
   }
   bool expect(const char *c) { return sym(c) || fail(std::string("expected '") + c + "'"); }

   bool value(Value &v) {
      const Token &t = peek();
      if (t.kind == T_NUMBER) {
         v.kind   = Value::NUMBER;
         v.number = t.number;
         pos++;
         return true;
      }
      if (t.kind == T_STRING) {
         v.kind = Value::STRING;
         v.text = t.text;
         pos++;
         return true;
      }
      if (t.kind == T_NULL) {
         v.kind = Value::NULLSTR;
         pos++;
         return true;
      }
      if (sym("[")) {
         v.kind = Value::ARRAY;
         if (sym("]")) {
            return true;
         }
         while (true) {
            if (peek().kind != T_NUMBER) {
               return fail("expected a number in an array");
            }
            v.array.push_back(peek().number);
            pos++;
            if (sym("]")) {
               return true;
            }
            if (!expect(",")) {
               return false;
            }
         }
      }
      return fail("expected a value");
   }

   bool group(Group &g) {
      g.keyword = peek().text;
      pos++;
      if (peek().kind != T_STRING) {
         return fail("expected the group name after " + g.keyword);
      }
      g.name = peek().text;
      pos++;
      if (!expect("=") || !expect("{")) {
         return false;
      }
      while (!sym("}")) {
         const Token &t = peek();
         if (t.kind == T_INCLUDE) {
            pos++;
            if (peek().kind != T_STRING) {
               return fail("expected a group name after #include");
            }
            std::string src = peek().text;
            pos++;
            sym(";");
            bool found = false;
            for (auto &o : groups) {
               if (o.name == src) {
                  for (auto &p : o.params) {
                     if (!g.find(p.key)) {
                        g.params.push_back(p);
                     }
                  }
                  found = true;
               }
            }
            if (!found) {
               return fail("#include of unknown group \"" + src + "\"");
            }
            continue;
         }
         if (t.kind != T_ID && t.kind != T_OVERWRITE) {
            return fail("expected a parameter name in group \"" + g.name + "\"");
         }
         bool over = t.kind == T_OVERWRITE;
         Param p;
         p.key = t.text;
         pos++;
         if (!expect("=") || !value(p.value) || !expect(";")) {
            return false;
         }
         Param *old = g.find(p.key);
         if (old) {
            if (!over) {
               return fail("parameter " + p.key + " defined twice in group \"" + g.name + "\"");
            }
            old->value = p.value;
         }
         else {
            if (over) {
               return fail("@" + p.key + " overwrites nothing in group \"" + g.name + "\"");
            }
            g.params.push_back(p);
         }
      }
      sym(";");
      return true;
   }

   bool parse() {
      while (peek().kind != T_END) {
         const Token &t = peek();
         if (t.kind != T_ID) {
            return fail("expected a group keyword or a directive");
         }
         if (t.text == "ParameterSweep" || t.text == "BatchSweep") {
            return fail(t.text + " is not supported");
         }
         if (pos + 1 < toks.size() && toks[pos + 1].kind == T_SYM && toks[pos + 1].text == "=") {
            Param d;
            d.key = t.text;
            pos += 2;
            if (!value(d.value) || !expect(";")) {
               return false;
            }
            directives.push_back(d);
            continue;
         }
         Group g;
         if (!group(g)) {
            return false;
         }
         for (auto &o : groups) {
            if (o.name == g.name) {
               return fail("group name \"" + g.name + "\" used twice");
            }
         }
         groups.push_back(g);
      }
      return true;
   }
};

} // namespace

struct pvparams {
   std::vector<Group> groups;
   std::vector<Param> directives;
};

extern "C" {

const char *pvparams_error(void) { return gError.c_str(); }

int pvparams_parse(const char *text, pvparams **out) {
   if (!text || !out) {
      gError = "NULL argument";
      return -1;
   }
   Parser ps;
   if (!tokenize(text, ps.toks) || !ps.parse()) {
      return -2;
   }
   auto *p       = new pvparams;
   p->groups     = std::move(ps.groups);
   p->directives = std::move(ps.directives);
   *out          = p;
   return 0;
}

int pvparams_load(const char *path, pvparams **out) {
   std::ifstream f(path);
   if (!f) {
      gError = std::string("cannot open ") + (path ? path : "(null)");
      return -1;
   }
   std::stringstream ss;
   ss << f.rdbuf();
   return pvparams_parse(ss.str().c_str(), out);
}

void pvparams_free(pvparams *p) { delete p; }

int pvparams_num_groups(const pvparams *p) { return p ? (int)p->groups.size() : 0; }

const char *pvparams_group_keyword(const pvparams *p, int g) {
   return (p && g >= 0 && g < (int)p->groups.size()) ? p->groups[g].keyword.c_str() : nullptr;
}

const char *pvparams_group_name(const pvparams *p, int g) {
   return (p && g >= 0 && g < (int)p->groups.size()) ? p->groups[g].name.c_str() : nullptr;
}

int pvparams_find_group(const pvparams *p, const char *name) {
   if (!p || !name) {
      return -1;
   }
   for (size_t i = 0; i < p->groups.size(); i++) {
      if (p->groups[i].name == name) {
         return (int)i;
      }
   }
   return -1;
}

static const Value *lookup(const pvparams *p, int g, const char *key) {
   if (!p || !key || g < 0 || g >= (int)p->groups.size()) {
      return nullptr;
   }
   for (auto &q : p->groups[g].params) {
      if (q.key == key) {
         return &q.value;
      }
   }
   return nullptr;
}

int pvparams_number(const pvparams *p, int g, const char *key, double *value) {
   const Value *v = lookup(p, g, key);
   if (!v) {
      return 0;
   }
   if (v->kind == Value::NUMBER) {
      *value = v->number;
      return 1;
   }
   if (v->kind == Value::ARRAY && !v->array.empty()) {
      *value = v->array[0];
      return 1;
   }
   return 0;
}

int pvparams_string(const pvparams *p, int g, const char *key, const char **value) {
   const Value *v = lookup(p, g, key);
   if (!v || (v->kind != Value::STRING && v->kind != Value::NULLSTR)) {
      return 0;
   }
   *value = v->kind == Value::NULLSTR ? nullptr : v->text.c_str();
   return 1;
}

int pvparams_array(const pvparams *p, int g, const char *key, double *values, int max) {
   const Value *v = lookup(p, g, key);
   if (!v) {
      return 0;
   }
   if (v->kind == Value::NUMBER) {
      if (max > 0 && values) {
         values[0] = v->number;
      }
      return 1;
   }
   if (v->kind != Value::ARRAY) {
      return 0;
   }
   for (int i = 0; i < (int)v->array.size() && i < max && values; i++) {
      values[i] = v->array[i];
   }
   return (int)v->array.size();
}

int pvparams_directive(const pvparams *p, const char *key, double *value) {
   if (!p || !key) {
      return 0;
   }
   for (auto &d : p->directives) {
      if (d.key == key && d.value.kind == Value::NUMBER) {
         *value = d.value.number;
         return 1;
      }
   }
   return 0;
}

} // extern "C"
