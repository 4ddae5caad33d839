// X3D shape reader: a streaming XML reader (declarations, comments, CDATA,
// doctype skipped; elements and attributes delivered as start/end events)
// driving the reference's shape handler rules (src/ShapeParser.cpp:79-147).
#include "x3d.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "bem_compat.h"

namespace {

[[noreturn]] void bad(const std::string &m) { throw bem::Error(BEM_E_INVALID, "x3d: " + m); }

// One transform of the stack (src/util/Transform.cpp:3-70).  A rotation is
// (axis x, y, z, angle phi) with s = sin(phi), c = cos(phi), t = 1 - c.
struct Rot {
  double x = 0, y = 0, z = 1, phi = 0;
  void rotate(const double v[3], double r[3], double sign) const {
    const double p = sign * phi, s = std::sin(p), c = std::cos(p), t = 1. - c;
    r[0] = (t * x * x + c) * v[0] + (t * x * y + s * z) * v[1] + (t * x * z - s * y) * v[2];
    r[1] = (t * x * y - s * z) * v[0] + (t * y * y + c) * v[1] + (t * y * z + s * x) * v[2];
    r[2] = (t * x * z + s * y) * v[0] + (t * y * z - s * x) * v[1] + (t * z * z + c) * v[2];
  }
};

struct Xform {
  double C[3] = {0, 0, 0}, S[3] = {1, 1, 1}, T[3] = {0, 0, 0};
  Rot R, SR;
  bool null = true;
  void apply(double p[3]) const {
    if (null) return;
    double r[3] = {p[0] - C[0], p[1] - C[1], p[2] - C[2]}, q[3];
    SR.rotate(r, q, -1.0);
    for (int i = 0; i < 3; i++) q[i] = S[i] * q[i];
    SR.rotate(q, r, 1.0);
    R.rotate(r, q, 1.0);
    for (int i = 0; i < 3; i++) p[i] = q[i] + C[i] + T[i];
  }
};

// The reference's number groups: skip spaces, read a number, skip to the
// next space; false where the text ends before the wanted count.
bool read_numbers(const char *s, int count, double *out, bool integer) {
  for (int k = 0; k < count; k++) {
    while (*s == ' ') s++;
    if (*s == 0) return false;
    out[k] = integer ? (double)std::atoi(s) : std::atof(s);
    if (k + 1 < count)
      while (*s && *s != ' ') s++;
  }
  return true;
}

void read_vector(const std::string &v, double out[3]) {
  if (!read_numbers(v.c_str(), 3, out, false)) bad("bad vector '" + v + "'");
}
void read_rotation(const std::string &v, Rot &r) {
  double q[4];
  if (!read_numbers(v.c_str(), 4, q, false)) bad("bad rotation '" + v + "'");
  r.x = q[0], r.y = q[1], r.z = q[2], r.phi = q[3];
}

struct Handler {
  X3DShape &shape;
  std::vector<Xform> stack;
  explicit Handler(X3DShape &s) : shape(s) {}

  static const std::string *find(const std::vector<std::pair<std::string, std::string>> &a, const char *k) {
    for (auto &kv : a)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }

  void start(const std::string &name, const std::vector<std::pair<std::string, std::string>> &attrs) {
    Xform t;
    if (name == "Transform") {
      if (auto v = find(attrs, "center")) {
        read_vector(*v, t.C);
        if (t.C[0] != 0. || t.C[1] != 0. || t.C[2] != 0.) t.null = false;
      }
      if (auto v = find(attrs, "rotation")) {
        read_rotation(*v, t.R);
        if (t.R.phi != 0.) t.null = false;
      }
      if (auto v = find(attrs, "scale")) {
        read_vector(*v, t.S);
        if (t.S[0] != 1. || t.S[1] != 1. || t.S[2] != 1.) t.null = false;
      }
      if (auto v = find(attrs, "scaleOrientation")) {
        read_rotation(*v, t.SR);
        if (t.SR.phi != 0.) t.null = false;
      }
      if (auto v = find(attrs, "translation")) {
        read_vector(*v, t.T);
        if (t.T[0] != 0. || t.T[1] != 0. || t.T[2] != 0.) t.null = false;
      }
    }
    stack.push_back(t);
    if (name == "Dielectric") {
      auto m = find(attrs, "material");
      if (!m) bad("<Dielectric> without a material attribute");
      shape.dielectric = true;
      shape.material = *m;
      shape.params.clear();
      for (auto &kv : attrs)
        if (kv.first != "material") shape.params.push_back(kv);
    } else if (name == "IndexedFaceSet") {
      if (auto v = find(attrs, "coordIndex")) {
        // each group is read from its start to the end of the attribute
        for (size_t i = 0;;) {
          double f[4];
          if (read_numbers(v->c_str() + i, 4, f, true))
            for (int k = 0; k < 3; k++) shape.face.push_back((int)f[k]);
          size_t c = v->find(',', i);
          if (c == std::string::npos) break;
          i = c + 1;
        }
      }
    } else if (name == "Coordinate") {
      if (auto v = find(attrs, "point")) {
        for (size_t i = 0;;) {
          double p[3];
          if (read_numbers(v->c_str() + i, 3, p, false)) {
            for (auto &x : stack) x.apply(p);
            shape.coord.insert(shape.coord.end(), p, p + 3);
          }
          size_t c = v->find(',', i);
          if (c == std::string::npos) break;
          i = c + 1;
        }
      }
    }
  }
  void end() {
    if (stack.empty()) bad("unbalanced end tag");
    stack.pop_back();
  }
};

bool name_char(char c) {
  return std::isalnum((unsigned char)c) || c == '_' || c == ':' || c == '-' || c == '.';
}

// Attribute value as an XML parser delivers it: literal tab, CR and LF become
// spaces, entities and character references are decoded.
std::string attr_value(const std::string &raw) {
  std::string out;
  for (size_t i = 0; i < raw.size(); i++) {
    char c = raw[i];
    if (c == '&') {
      size_t e = raw.find(';', i);
      if (e == std::string::npos) bad("unterminated entity");
      std::string ent = raw.substr(i + 1, e - i - 1);
      if (ent == "lt") out += '<';
      else if (ent == "gt") out += '>';
      else if (ent == "amp") out += '&';
      else if (ent == "quot") out += '"';
      else if (ent == "apos") out += '\'';
      else if (!ent.empty() && ent[0] == '#') {
        long code = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                                                         : std::strtol(ent.c_str() + 1, nullptr, 10);
        if (code <= 0 || code > 127) bad("character reference outside ASCII: &" + ent + ";");
        out += (char)code;
      } else {
        bad("unknown entity &" + ent + ";");
      }
      i = e;
    } else {
      out += (c == '\t' || c == '\n' || c == '\r') ? ' ' : c;
    }
  }
  return out;
}

void parse(const std::string &x, Handler &h) {
  size_t i = 0;
  const size_t n = x.size();
  auto skip_past = [&](const char *end) {
    size_t e = x.find(end, i);
    if (e == std::string::npos) bad(std::string("missing '") + end + "'");
    i = e + std::strlen(end);
  };
  int depth = 0;
  while (i < n) {
    if (x[i] != '<') {
      i++;
      continue;
    }
    if (x.compare(i, 4, "<!--") == 0) {
      skip_past("-->");
    } else if (x.compare(i, 9, "<![CDATA[") == 0) {
      skip_past("]]>");
    } else if (x.compare(i, 2, "<?") == 0) {
      skip_past("?>");
    } else if (x.compare(i, 2, "<!") == 0) {   // doctype, with an optional internal subset
      int br = 0;
      for (i += 2; i < n; i++) {
        if (x[i] == '[') br++;
        else if (x[i] == ']') br--;
        else if (x[i] == '>' && br <= 0) break;
      }
      if (i >= n) bad("unterminated declaration");
      i++;
    } else if (x.compare(i, 2, "</") == 0) {
      skip_past(">");
      if (--depth < 0) bad("unbalanced end tag");
      h.end();
    } else {
      size_t j = i + 1;
      while (j < n && name_char(x[j])) j++;
      std::string name = x.substr(i + 1, j - i - 1);
      if (name.empty()) bad("bad tag");
      std::vector<std::pair<std::string, std::string>> attrs;
      bool closed = false;
      for (;;) {
        while (j < n && std::isspace((unsigned char)x[j])) j++;
        if (j >= n) bad("unterminated tag <" + name);
        if (x[j] == '>') {
          j++;
          break;
        }
        if (x[j] == '/' && j + 1 < n && x[j + 1] == '>') {
          j += 2;
          closed = true;
          break;
        }
        size_t k = j;
        while (k < n && name_char(x[k])) k++;
        if (k == j) bad("bad attribute in <" + name + ">");
        std::string key = x.substr(j, k - j);
        while (k < n && std::isspace((unsigned char)x[k])) k++;
        if (k >= n || x[k] != '=') bad("attribute without value in <" + name + ">");
        k++;
        while (k < n && std::isspace((unsigned char)x[k])) k++;
        if (k >= n || (x[k] != '"' && x[k] != '\'')) bad("unquoted attribute in <" + name + ">");
        const char q = x[k];
        size_t e = x.find(q, k + 1);
        if (e == std::string::npos) bad("unterminated attribute in <" + name + ">");
        attrs.emplace_back(key, attr_value(x.substr(k + 1, e - k - 1)));
        j = e + 1;
      }
      i = j;
      h.start(name, attrs);
      if (closed) h.end();
      else depth++;
    }
  }
  if (depth != 0) bad("unclosed elements at end of document");
}

}  // namespace

void read_x3d_text(const std::string &text, X3DShape &shape) {
  Handler h(shape);
  parse(text, h);
}

void read_x3d(const char *path, X3DShape &shape) {
  std::ifstream f(path, std::ios::binary);
  if (!f) bad(std::string("cannot open ") + path);
  std::stringstream ss;
  ss << f.rdbuf();
  read_x3d_text(ss.str(), shape);
}

// C entry points (declared in x3d.h).
extern "C" {
int bemc_x3d_read(const char *path, double *xyz, int *nv, int *tri, int *nt, char *material, int matcap) {
  try {
    X3DShape s;
    read_x3d(path, s);
    const int v = (int)(s.coord.size() / 3), t = (int)(s.face.size() / 3);
    if (xyz && *nv < v) return BEM_E_INVALID;
    if (tri && *nt < t) return BEM_E_INVALID;
    if (xyz) std::copy(s.coord.begin(), s.coord.end(), xyz);
    if (tri) std::copy(s.face.begin(), s.face.end(), tri);
    *nv = v;
    *nt = t;
    if (material && matcap > 0) {
      std::snprintf(material, (size_t)matcap, "%s", s.material.c_str());
    }
    return 0;
  } catch (const bem::Error &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return e.code;
  }
}

int bemc_dielectric(const char *name, const char *params, double eV, double T, double *epsmu4) {
  try {
    Params p;
    if (params) {
      std::string all(params);
      for (size_t i = 0; i < all.size();) {
        size_t e = all.find(';', i);
        if (e == std::string::npos) e = all.size();
        std::string kv = all.substr(i, e - i);
        size_t q = kv.find('=');
        if (q == std::string::npos) return BEM_E_INVALID;
        p.params.emplace_back(kv.substr(0, q), kv.substr(q + 1));
        i = e + 1;
      }
    }
    Dielectric *d = Dielectrics::getDielectric(name, p);
    d->setTemperature(T);
    const Complex e = d->eps(eV), m = d->mu(eV);
    epsmu4[0] = e.real(), epsmu4[1] = e.imag(), epsmu4[2] = m.real(), epsmu4[3] = m.imag();
    return 0;
  } catch (const bem::Error &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return e.code;
  }
}
}
