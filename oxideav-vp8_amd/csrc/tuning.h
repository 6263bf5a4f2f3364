// Launch-shape tuning table (host only).
#pragma once
#include <string>
#include <vector>

namespace vp8g {

struct TuneEntry {
  char cls[16];    // resolution class or "batch"
  int n;           // group length (0 for "batch")
  int mbs, waves, wgs_per_cu;
};

// qcif, cif, 360p, 720p, 1080p or 4k by luma area.
const char* resolution_class(int w, int h);
bool parse_tuning(const std::string& text, std::vector<TuneEntry>* out);
bool load_tuning_file(const char* path, std::vector<TuneEntry>* out);
// Entry of class `cls` with the group length nearest `n` (the shorter on a
// tie), or nullptr.
const TuneEntry* lookup_tuning(const std::vector<TuneEntry>& t, const char* cls, int n);

}  // namespace vp8g
