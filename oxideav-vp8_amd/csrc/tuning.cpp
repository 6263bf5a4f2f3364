// Launch-shape tuning table: the JSON file the autotuner writes, read once
// per context.  Only the subset the autotuner emits is understood: a list
// of flat objects with string and number values; every object that has a
// "class" key is an entry.  Unknown keys are ignored.
#include "tuning.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

namespace vp8g {

namespace {

// Value text of "key" inside one flat object, or npos.
size_t find_value(const std::string& obj, const char* key) {
  const std::string k = std::string("\"") + key + "\"";
  size_t p = obj.find(k);
  if (p == std::string::npos) return p;
  p = obj.find(':', p + k.size());
  if (p == std::string::npos) return p;
  p++;
  while (p < obj.size() && (obj[p] == ' ' || obj[p] == '\t' || obj[p] == '\n' || obj[p] == '\r')) p++;
  return p < obj.size() ? p : std::string::npos;
}

bool get_str(const std::string& obj, const char* key, char* out, size_t cap) {
  size_t p = find_value(obj, key);
  if (p == std::string::npos || obj[p] != '"') return false;
  const size_t e = obj.find('"', p + 1);
  if (e == std::string::npos || e - p - 1 >= cap) return false;
  memcpy(out, obj.data() + p + 1, e - p - 1);
  out[e - p - 1] = 0;
  return true;
}

bool get_int(const std::string& obj, const char* key, int* out) {
  size_t p = find_value(obj, key);
  if (p == std::string::npos) return false;
  char* end = nullptr;
  const long v = strtol(obj.c_str() + p, &end, 10);
  if (end == obj.c_str() + p || v < -1000000 || v > 1000000) return false;
  *out = (int)v;
  return true;
}

}  // namespace

const char* resolution_class(int w, int h) {
  const long long a = (long long)w * h;
  if (a <= 176LL * 144) return "qcif";
  if (a <= 352LL * 288) return "cif";
  if (a <= 640LL * 360) return "360p";
  if (a <= 1280LL * 720) return "720p";
  if (a <= 1920LL * 1088) return "1080p";
  return "4k";
}

bool parse_tuning(const std::string& text, std::vector<TuneEntry>* out) {
  std::vector<TuneEntry> entries;
  size_t p = 0;
  while ((p = text.find('{', p)) != std::string::npos) {
    const size_t close = text.find('}', p + 1);
    if (close == std::string::npos) return false;
    const size_t next_open = text.find('{', p + 1);
    if (next_open != std::string::npos && next_open < close) {  // not innermost
      p = next_open;
      continue;
    }
    const std::string obj = text.substr(p, close - p + 1);
    p = close + 1;
    TuneEntry e;
    if (!get_str(obj, "class", e.cls, sizeof(e.cls))) continue;
    if (!get_int(obj, "n", &e.n)) e.n = 0;
    if (!get_int(obj, "mbs", &e.mbs) || !get_int(obj, "waves", &e.waves)) return false;
    if (!get_int(obj, "wgs_per_cu", &e.wgs_per_cu)) e.wgs_per_cu = 0;
    if (e.mbs < 0 || e.mbs > 8 || e.waves < 0 || e.waves > 16 || e.wgs_per_cu < 0 || e.wgs_per_cu > 32)
      return false;
    entries.push_back(e);
  }
  *out = entries;
  return true;
}

bool load_tuning_file(const char* path, std::vector<TuneEntry>* out) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  std::string text;
  char buf[4096];
  size_t got;
  while ((got = fread(buf, 1, sizeof(buf), f)) > 0) {
    text.append(buf, got);
    if (text.size() > (1u << 20)) {
      fclose(f);
      return false;
    }
  }
  fclose(f);
  return parse_tuning(text, out);
}

const TuneEntry* lookup_tuning(const std::vector<TuneEntry>& t, const char* cls, int n) {
  const TuneEntry* best = nullptr;
  for (const TuneEntry& e : t) {
    if (strcmp(e.cls, cls) != 0) continue;
    if (!best) {
      best = &e;
      continue;
    }
    const int d = abs(e.n - n), bd = abs(best->n - n);
    if (d < bd || (d == bd && e.n < best->n)) best = &e;
  }
  return best;
}

}  // namespace vp8g
