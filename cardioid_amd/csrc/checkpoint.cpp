// checkpoint.cpp -- the reference's checkpoint / restart files for this path.
//
// Writer (elec/checkpointIO.cc:56-207): directory snapshot.<loop, 12
// digits>, one pio file "state#000000" whose header (writeHeader, :56-100)
// is a `state FILEHEADER` object followed by a blank line, then one
// fixed-length record per cell: gid, Vm and the reaction's checkpoint
// fields in the order ReactionManager::getCheckpointInfo lists them (a
// std::map, so sorted by name), as "%12llu %21.13e" + " %21.13e" per field
// + "\n" (FIXRECORDASCII, lrec = 34 + 22 nfields + 1) or as 8-byte words
// (FIXRECORDBINARY, lrec = 8 (nfields + 2)); a `restart` file holding the
// SIMULATE object's loop, time and stateFile (writeRestart, :103-119); and
// a `restart` symlink beside the snapshot directory (:200-205).
//
// Reader (readCheckpoint, :209-268, with stateLoader.cc:92-127 and
// simUtil/src/pio.c:428-520, 642-700): the header is every line up to and
// including the first one of at most one character; the records follow it
// in file #000000 and fill files #000001.. in turn; the gid column is found
// by name; fields are matched by name through the reaction's handles and
// scaled by the unit conversion to the model's unit; fields the file lacks
// keep their values; Vm is taken as written.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <unordered_map>
#include <vector>

#include "../../include/cardioid_hip.h"
#include "hip_util.hpp"
#include "tt06_model.hpp"

using namespace cdh;

namespace {

struct Field {
   std::string name, unit;
   int slot;
};

// the 19 TT06Dev fields sorted by name, as std::map orders them
std::vector<Field> sorted_fields()
{
   std::vector<Field> f;
   for (int i = 0; i < CDH_NSTATE; i++) {
      const char *name = nullptr, *unit = nullptr;
      int handle = 0;
      cdh_tt06_checkpoint_field(i, &name, &unit, &handle);
      f.push_back({name, unit, handle});
   }
   std::sort(f.begin(), f.end(), [](const Field &a, const Field &b) { return a.name < b.name; });
   return f;
}

std::string join(const std::vector<std::string> &v)
{
   std::string s;
   for (size_t i = 0; i < v.size(); i++) s += (i ? " " : "") + v[i];
   return s;
}

void make_dir(const std::string &d)
{
   if (mkdir(d.c_str(), 0775) != 0 && errno != EEXIST)
      throw Error(kErrArg, "checkpoint: cannot create directory " + d);
}

std::string trim(std::string s)
{
   const char *ws = " \t\n\r";
   s.erase(0, s.find_first_not_of(ws));
   const size_t e = s.find_last_not_of(ws);
   s.erase(e == std::string::npos ? 0 : e + 1);
   return s;
}

// key -> value of a `name TYPE { key = value; ... }` object
std::map<std::string, std::string> parse_object(const std::string &text)
{
   std::map<std::string, std::string> kv;
   const size_t open = text.find('{'), close = text.rfind('}');
   if (open == std::string::npos || close == std::string::npos || close < open)
      throw Error(kErrArg, "checkpoint: no FILEHEADER object");
   std::stringstream body(text.substr(open + 1, close - open - 1));
   std::string item;
   while (std::getline(body, item, ';')) {
      const size_t eq = item.find('=');
      if (eq == std::string::npos) continue;
      kv[trim(item.substr(0, eq))] = trim(item.substr(eq + 1));
   }
   return kv;
}

std::vector<std::string> words(const std::string &s)
{
   std::vector<std::string> w;
   std::stringstream ss(s);
   std::string x;
   while (ss >> x) w.push_back(x);
   return w;
}

// multiplier taking a value in `from` to `to` for the units this model's
// fields use (dimensionless, voltage, concentration); equal names need none
double unit_factor(const std::string &from, const std::string &to)
{
   if (from == to) return 1.0;
   struct U {
      const char *name;
      int dim;
      double si;
   };
   static const U table[] = {{"1", 0, 1.0},   {"V", 1, 1.0}, {"mV", 1, 1e-3}, {"uV", 1, 1e-6},
                             {"M", 2, 1.0},   {"mM", 2, 1e-3}, {"uM", 2, 1e-6},
                             {"nM", 2, 1e-9}};
   const U *a = nullptr, *b = nullptr;
   for (const U &u : table) {
      if (from == u.name) a = &u;
      if (to == u.name) b = &u;
   }
   CDH_REQUIRE(a && b && a->dim == b->dim,
               "checkpoint: cannot convert from unit " + from + " to " + to);
   return a->si / b->si;
}

// "#%6.6d" file name of pio file `fid` of `base`
std::string pio_name(const std::string &base, int fid)
{
   char num[16];
   std::snprintf(num, sizeof num, "%06d", fid);
   return base + "#" + num;
}

}  // namespace

extern "C" {

int cdh_checkpoint_write(const char *dir, const char *simulate_name, int loop, double time, int nx,
                         int ny, int nz, int64_t n, const int64_t *gid, const double *Vm,
                         const double *state, int ascii, char *state_path, int path_len)
{
   return guard([&] {
      CDH_REQUIRE(dir && n >= 0 && (n == 0 || (gid && Vm && state)),
                  "cdh_checkpoint_write: null argument");
      const std::vector<Field> fields = sorted_fields();
      char num[32];
      std::snprintf(num, sizeof num, "snapshot.%012d", loop);
      const std::string snap = num;                  // relative to dir, as in the restart file
      const std::string dirName = std::string(dir) + "/" + snap;
      make_dir(dirName);
      const std::string stateRel = snap + "/state";
      const std::string stateBase = std::string(dir) + "/" + stateRel;

      std::vector<std::string> names, types, units;
      for (const Field &f : fields) {
         names.push_back(f.name);
         units.push_back(f.unit);
         types.push_back(ascii ? "f" : "f8");
      }
      const int nf = (int)fields.size();
      const int lrec = ascii ? 34 + 22 * nf + 1 : 8 * (nf + 2);

      const std::string fileName = pio_name(stateBase, 0);
      std::FILE *fp = std::fopen(fileName.c_str(), "wb");
      CDH_REQUIRE(fp, "cdh_checkpoint_write: cannot open " + fileName);
      int endianKey;
      std::memcpy(&endianKey, "1234", 4);
      char host[256] = "unknown";
      gethostname(host, sizeof host - 1);
      const char *user = std::getenv("USER");
      const std::time_t now = std::time(nullptr);
      const std::string stamp = trim(std::asctime(std::localtime(&now)));
      std::fprintf(fp, "state FILEHEADER {\n");
      std::fprintf(fp, "   create_time = %s;\n", stamp.c_str());
      std::fprintf(fp, "   user = %s;", user ? user : "unknown");
      std::fprintf(fp, "   host = %s;\n", host);
      std::fprintf(fp, "   exe_version = cardioid_amd;");
      std::fprintf(fp, "   srcpath = cardioid_amd;\n");
      std::fprintf(fp, "   compile_time = %s %s;\n", __DATE__, __TIME__);
      std::fprintf(fp, "   datatype = %s;\n", ascii ? "FIXRECORDASCII" : "FIXRECORDBINARY");
      std::fprintf(fp, "   nfiles = %d;\n", 1);
      std::fprintf(fp, "   nrecord = %lld;\n", (long long)n);
      std::fprintf(fp, "   lrec = %d;\n", lrec);
      std::fprintf(fp, "   endian_key = %d;\n", endianKey);
      std::fprintf(fp, "   time = %f;\n", time);
      std::fprintf(fp, "   loop = %d;\n", loop);
      std::fprintf(fp, "   reactionMethod = %s;\n", "TT06Dev");
      std::fprintf(fp, "   nfields = %d;\n", nf + 2);
      std::fprintf(fp, "   field_names = gid Vm %s;\n", join(names).c_str());
      std::fprintf(fp, "   field_types = %s %s;\n", ascii ? "u f" : "u8 f8", join(types).c_str());
      std::fprintf(fp, "   field_units = 1 mV %s;\n", join(units).c_str());
      std::fprintf(fp, "   nx = %u; ny = %u; nz = %u;\n", (unsigned)nx, (unsigned)ny, (unsigned)nz);
      std::fprintf(fp, "}\n\n");
      std::vector<char> buf(lrec + 64);
      bool ok = true;
      for (int64_t c = 0; c < n && ok; c++) {
         if (ascii) {
            int pos = std::snprintf(buf.data(), buf.size(), "%12llu %21.13e",
                                    (unsigned long long)gid[c], Vm[c]);
            for (const Field &f : fields)
               if (pos < (int)buf.size())
                  pos += std::snprintf(buf.data() + pos, buf.size() - pos, " %21.13e",
                                       state[(int64_t)f.slot * n + c]);
            ok = pos + 1 == lrec;
            if (ok) buf[pos] = '\n';
         } else {
            std::memcpy(buf.data(), &gid[c], 8);
            std::memcpy(buf.data() + 8, &Vm[c], 8);
            for (int i = 0; i < nf; i++)
               std::memcpy(buf.data() + 16 + 8 * i, &state[(int64_t)fields[i].slot * n + c], 8);
         }
         ok = ok && std::fwrite(buf.data(), 1, lrec, fp) == (size_t)lrec;
      }
      const bool closed = std::fclose(fp) == 0;
      CDH_REQUIRE(ok, "cdh_checkpoint_write: a record does not fit its length (gid above 12 "
                      "digits?) or the write failed");
      CDH_REQUIRE(closed, "cdh_checkpoint_write: close failed");

      const std::string restart = dirName + "/restart";
      std::FILE *rf = std::fopen(restart.c_str(), "w");
      CDH_REQUIRE(rf, "cdh_checkpoint_write: cannot open " + restart);
      std::fprintf(rf, "%s SIMULATE {\n   loop=%d; time=%f; stateFile=%s;\n}\n",
                   simulate_name ? simulate_name : "simulate", loop, time, stateRel.c_str());
      CDH_REQUIRE(std::fclose(rf) == 0, "cdh_checkpoint_write: close failed");
      const std::string link = std::string(dir) + "/restart";
      unlink(link.c_str());
      CDH_REQUIRE(symlink((snap + "/restart").c_str(), link.c_str()) == 0,
                  "cdh_checkpoint_write: cannot link " + link);
      if (state_path && path_len > 0) {
         std::strncpy(state_path, stateBase.c_str(), path_len - 1);
         state_path[path_len - 1] = '\0';
      }
   });
}

int cdh_checkpoint_read(const char *state_file, int64_t n, const int64_t *gid, double *Vm,
                        double *state, int *loop, double *time)
{
   return guard([&] {
      CDH_REQUIRE(state_file && n >= 0 && (n == 0 || gid), "cdh_checkpoint_read: null argument");
      // accepts "<base>", "<base>#" or "<base>#000000"
      std::string base = state_file;
      const size_t hash = base.rfind('#');
      if (hash != std::string::npos && base.size() - hash <= 7) base.erase(hash);
      std::ifstream in(pio_name(base, 0), std::ios::binary);
      CDH_REQUIRE(in.good(), "cdh_checkpoint_read: cannot open " + pio_name(base, 0));
      // header: every line up to and including the first of at most one
      // character (the blank line after the object)
      std::string header, line;
      for (;;) {
         CDH_REQUIRE((bool)std::getline(in, line), "cdh_checkpoint_read: header not terminated");
         header += line + "\n";
         if (line.empty()) break;
      }
      const auto kv = parse_object(header);
      auto get = [&](const char *k) {
         auto it = kv.find(k);
         CDH_REQUIRE(it != kv.end(), std::string("cdh_checkpoint_read: header has no ") + k);
         return it->second;
      };
      const std::string datatype = get("datatype");
      const bool ascii = datatype == "FIXRECORDASCII";
      CDH_REQUIRE(ascii || datatype == "FIXRECORDBINARY",
                  "cdh_checkpoint_read: unsupported datatype " + datatype);
      const int nfiles = std::stoi(get("nfiles"));
      const int64_t nrec = std::stoll(get("nrecord"));
      const int lrec = std::stoi(get("lrec"));
      CDH_REQUIRE(nfiles >= 1 && lrec > 0 && nrec >= 0, "cdh_checkpoint_read: bad header");
      const std::vector<std::string> names = words(get("field_names"));
      const std::vector<std::string> types = words(get("field_types"));
      const std::vector<std::string> units = words(get("field_units"));
      CDH_REQUIRE(names.size() == types.size() && names.size() == units.size(),
                  "cdh_checkpoint_read: field_names, field_types, field_units disagree");
      const auto gidIt = std::find(names.begin(), names.end(), "gid");
      CDH_REQUIRE(gidIt != names.end(), "cdh_checkpoint_read: no gid field");
      const size_t gidCol = gidIt - names.begin();
      CDH_REQUIRE(nrec == n, "cdh_checkpoint_read: " + std::to_string(nrec) +
                                 " records for " + std::to_string(n) + " cells");
      if (loop) *loop = std::stoi(get("loop"));
      if (time) *time = std::stod(get("time"));
      bool swap = false;
      if (!ascii) {
         int key;
         std::memcpy(&key, "1234", 4);
         const long long fileKey = std::stoll(get("endian_key"));
         const int swapped = (int)__builtin_bswap32((uint32_t)key);
         CDH_REQUIRE(fileKey == key || fileKey == swapped, "cdh_checkpoint_read: bad endian_key");
         swap = fileKey != key;
      }

      // per column: byte offset and width (binary), integer or not; target
      // -2 Vm, >= 0 a TT06Dev slot, -1 ignored
      const size_t nc = names.size();
      std::vector<int> target(nc, -1), width(nc, 0), offset(nc, 0);
      std::vector<char> isInt(nc, 0);
      std::vector<double> scale(nc, 1.0);
      int pos = 0;
      for (size_t k = 0; k < nc; k++) {
         const std::string &t = types[k];
         isInt[k] = t == "u" || t == "d" || t == "u8";
         if (!ascii) {
            CDH_REQUIRE(t == "u8" || t == "f8" || t == "f4",
                        "cdh_checkpoint_read: binary field type " + t);
            width[k] = t == "f4" ? 4 : 8;
            offset[k] = pos;
            pos += width[k];
         } else {
            CDH_REQUIRE(t == "u" || t == "d" || t == "f" || t == "s",
                        "cdh_checkpoint_read: ASCII field type " + t);
         }
         if (k == gidCol) {
            CDH_REQUIRE(isInt[k], "cdh_checkpoint_read: gid is not an integer field");
            continue;
         }
         if (names[k] == "Vm") {
            target[k] = -2;
            continue;
         }
         const int h = cdh_tt06_var_handle(names[k].c_str());
         if (h < 0) continue;
         CDH_REQUIRE(t != "s", "cdh_checkpoint_read: string field " + names[k]);
         const char *nm = nullptr, *un = nullptr;
         int hh = 0;
         cdh_tt06_checkpoint_field(h, &nm, &un, &hh);
         target[k] = h;
         scale[k] = unit_factor(units[k], un);
      }
      CDH_REQUIRE(ascii || pos <= lrec, "cdh_checkpoint_read: lrec shorter than the fields");

      std::unordered_map<int64_t, int64_t> where;
      where.reserve(n * 2);
      for (int64_t c = 0; c < n; c++)
         CDH_REQUIRE(where.emplace(gid[c], c).second, "cdh_checkpoint_read: duplicate cell gid");
      std::vector<char> buf(lrec + 1, 0);
      std::vector<char> seen(n, 0);
      std::vector<std::string> tok(nc);
      std::vector<double> val(nc, 0.0);
      int fid = 0;
      for (int64_t r = 0; r < nrec; r++) {
         in.read(buf.data(), lrec);
         while (in.gcount() == 0 && fid + 1 < nfiles) {
            in.close();
            in.clear();
            in.open(pio_name(base, ++fid), std::ios::binary);
            CDH_REQUIRE(in.good(), "cdh_checkpoint_read: cannot open " + pio_name(base, fid));
            in.read(buf.data(), lrec);
         }
         CDH_REQUIRE(in.gcount() == lrec, "cdh_checkpoint_read: records end early");
         int64_t g = 0;
         if (ascii) {
            buf[lrec] = '\0';
            std::stringstream ss(buf.data());
            for (size_t k = 0; k < nc; k++) ss >> tok[k];
            CDH_REQUIRE(!ss.fail(), "cdh_checkpoint_read: record " + std::to_string(r) +
                                        " has too few fields");
            for (size_t k = 0; k < nc; k++) {
               if (k != gidCol && target[k] == -1) continue;
               char *end = nullptr;
               if (isInt[k]) {
                  const long long x = std::strtoll(tok[k].c_str(), &end, 10);
                  if (k == gidCol) g = x;
                  val[k] = (double)x;
               } else {
                  val[k] = std::strtod(tok[k].c_str(), &end);
               }
               CDH_REQUIRE(end && *end == '\0', "cdh_checkpoint_read: unreadable value " + tok[k]);
            }
         } else {
            for (size_t k = 0; k < nc; k++) {
               unsigned char b[8];
               std::memcpy(b, buf.data() + offset[k], width[k]);
               if (swap) std::reverse(b, b + width[k]);
               if (width[k] == 4) {
                  float f;
                  std::memcpy(&f, b, 4);
                  val[k] = f;
               } else if (isInt[k]) {
                  int64_t x;
                  std::memcpy(&x, b, 8);
                  if (k == gidCol) g = x;
                  val[k] = (double)x;
               } else {
                  std::memcpy(&val[k], b, 8);
               }
            }
         }
         auto it = where.find(g);
         CDH_REQUIRE(it != where.end(), "cdh_checkpoint_read: record for gid " +
                                            std::to_string(g) + ", which no cell has");
         const int64_t c = it->second;
         CDH_REQUIRE(!seen[c], "cdh_checkpoint_read: two records for gid " + std::to_string(g));
         seen[c] = 1;
         for (size_t k = 0; k < nc; k++) {
            if (target[k] == -2 && Vm) Vm[c] = val[k];
            else if (target[k] >= 0 && state)
               state[(int64_t)target[k] * n + c] = val[k] * scale[k];
         }
      }
   });
}

}  // extern "C"
