// demon_fits.cpp -- FITS writer/reader for DEMON's output file (see
// demon_fits.h).  FITS layout (FITS Standard 4.0): 2880-byte blocks, headers
// of 80-character cards padded with spaces, binary-table data as big-endian
// IEEE doubles padded with zeros.
#include "demon_fits.h"

#include <cctype>
#include <cmath>
#include <cstring>
#include <map>

#include "demon_host.h"

namespace demon_amd {

namespace {

constexpr long kBlock = 2880;
constexpr int kCard = 80;

long padded(long bytes) { return (bytes + kBlock - 1) / kBlock * kBlock; }

// One 80-character card.
std::string card(const std::string &text) {
    std::string c = text.substr(0, kCard);
    c.resize(kCard, ' ');
    return c;
}

// Fixed-format value card: keyword in columns 1-8, "= " in 9-10, the value
// right-justified to column 30 (numbers, logicals) or a quoted string from
// column 11 padded to at least 8 characters, then " / comment".
std::string value_card(const std::string &key, const std::string &value, const std::string &comment,
                       bool is_string = false) {
    std::string k = key;
    k.resize(8, ' ');
    std::string v;
    if (is_string) {
        std::string s = value;
        if (s.size() < 8) s.resize(8, ' ');
        v = "'" + s + "'";
        if (v.size() < 20) v.resize(20, ' ');
    } else {
        v = std::string(value.size() < 20 ? 20 - value.size() : 0, ' ') + value;
    }
    return card(k + "= " + v + (comment.empty() ? "" : " / " + comment));
}

std::string int_card(const std::string &key, long long v, const std::string &comment) {
    return value_card(key, std::to_string(v), comment);
}

std::string upper(std::string s) {
    for (char &ch : s) ch = (char)std::toupper((unsigned char)ch);
    return s;
}

// fits_write_key_dbl with 6 decimals: the value as %.6E; the keyword in upper
// case, as a HIERARCH card when longer than 8 characters (keywords match
// without regard to case).
std::string dbl_card(const std::string &key, double v, const std::string &comment) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.6E", v);
    if (key.size() <= 8) return value_card(upper(key), buf, comment);
    return card("HIERARCH " + upper(key) + " = " + buf + (comment.empty() ? "" : " / " + comment));
}

void put_be(unsigned char *dst, double v) {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    for (int k = 7; k >= 0; --k) {
        dst[k] = (unsigned char)(b & 0xff);
        b >>= 8;
    }
}

double get_be(const unsigned char *src) {
    uint64_t b = 0;
    for (int k = 0; k < 8; ++k) b = (b << 8) | src[k];
    double v;
    std::memcpy(&v, &b, 8);
    return v;
}

void write_all(std::FILE *f, const void *p, size_t n, const std::string &what) {
    if (n && std::fwrite(p, 1, n, f) != n) throw Error(DEMON_ERR_ARG, "FITS write failed: " + what);
}

// Header cards -> padded header bytes.
std::string header_bytes(const std::vector<std::string> &cards) {
    std::string h;
    for (const std::string &c : cards) h += c;
    h += card("END");
    h.resize((size_t)padded((long)h.size()), ' ');
    return h;
}

std::vector<std::string> table_header(long width, long rows, const std::vector<std::string> &type,
                                      const std::vector<std::string> &form, const std::vector<std::string> &unit,
                                      const std::string &name) {
    std::vector<std::string> c;
    c.push_back(value_card("XTENSION", "BINTABLE", "binary table extension", true));
    c.push_back(int_card("BITPIX", 8, "8-bit bytes"));
    c.push_back(int_card("NAXIS", 2, "2-dimensional binary table"));
    c.push_back(int_card("NAXIS1", width, "width of table in bytes"));
    c.push_back(int_card("NAXIS2", rows, "number of rows in table"));
    c.push_back(int_card("PCOUNT", 0, "size of special data area"));
    c.push_back(int_card("GCOUNT", 1, "one data group (required keyword)"));
    c.push_back(int_card("TFIELDS", (long long)type.size(), "number of fields in each row"));
    for (size_t k = 0; k < type.size(); ++k) {
        const std::string i = std::to_string(k + 1);
        c.push_back(value_card("TTYPE" + i, type[k], "label for field " + i, true));
        c.push_back(value_card("TFORM" + i, form[k], "data format of field: 8-byte DOUBLE", true));
        c.push_back(value_card("TUNIT" + i, unit[k], "physical unit of field", true));
    }
    c.push_back(value_card("EXTNAME", name, "name of this binary table extension", true));
    return c;
}

// One parsed header: keyword -> raw value text, plus where it lies.
struct Hdu {
    std::map<std::string, std::string> kv;   // standard keywords (upper case) and HIERARCH names
    std::vector<std::string> order;          // keyword order (for the primary header's keys)
    std::map<std::string, std::string> comment;
    std::map<std::string, long> card_at;     // byte offset of each card
    long header_start = 0, data_start = 0, data_bytes = 0;
};

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// Splits "value / comment" (a quoted string may hold '/').
void split_value(const std::string &rest, std::string &value, std::string &comment) {
    std::string r = trim(rest);
    size_t cut = std::string::npos;
    if (!r.empty() && r[0] == '\'') {
        size_t k = 1;
        for (; k < r.size(); ++k) {
            if (r[k] == '\'') {
                if (k + 1 < r.size() && r[k + 1] == '\'') ++k;
                else break;
            }
        }
        cut = r.find('/', k);
        value = trim(r.substr(1, k - 1));
    } else {
        cut = r.find('/');
        value = trim(r.substr(0, cut));
    }
    comment = cut == std::string::npos ? "" : trim(r.substr(cut + 1));
}

long long as_int(const Hdu &h, const std::string &k) {
    auto it = h.kv.find(k);
    if (it == h.kv.end()) throw Error(DEMON_ERR_ARG, "FITS header lacks " + k);
    return std::stoll(it->second);
}

// Reads the HDU starting at `at`; returns false at the end of the file.
bool read_hdu(std::FILE *f, long at, Hdu &h) {
    h = Hdu{};
    h.header_start = at;
    if (std::fseek(f, at, SEEK_SET)) return false;
    char buf[kCard];
    long pos = at;
    bool end = false;
    while (!end) {
        if (std::fread(buf, 1, kCard, f) != (size_t)kCard) {
            if (pos != at) throw Error(DEMON_ERR_ARG, "truncated FITS header");
            return false;
        }
        const std::string c(buf, kCard);
        const std::string key8 = trim(c.substr(0, 8));
        if (key8 == "END") {
            end = true;
        } else if (key8 == "HIERARCH") {
            const size_t eq = c.find('=');
            if (eq != std::string::npos) {
                const std::string name = trim(c.substr(8, eq - 8));
                std::string v, cm;
                split_value(c.substr(eq + 1), v, cm);
                h.kv[name] = v;
                h.comment[name] = cm;
                h.order.push_back(name);
                h.card_at[name] = pos;
            }
        } else if (!key8.empty() && c.compare(8, 2, "= ") == 0) {
            std::string v, cm;
            split_value(c.substr(10), v, cm);
            h.kv[key8] = v;
            h.comment[key8] = cm;
            h.order.push_back(key8);
            h.card_at[key8] = pos;
        }
        pos += kCard;
    }
    h.data_start = padded(pos);
    const long long naxis = as_int(h, "NAXIS");
    long long bytes = naxis > 0 ? std::llabs(as_int(h, "BITPIX")) / 8 : 0;
    for (long long k = 1; k <= naxis; ++k) bytes *= as_int(h, "NAXIS" + std::to_string(k));
    if (h.kv.count("PCOUNT")) bytes += as_int(h, "PCOUNT");
    h.data_bytes = (long)bytes;
    return true;
}

std::vector<Hdu> read_all(std::FILE *f) {
    std::vector<Hdu> out;
    long at = 0;
    Hdu h;
    while (read_hdu(f, at, h)) {
        out.push_back(h);
        at = h.data_start + padded(h.data_bytes);
    }
    if (out.empty() || out[0].kv["SIMPLE"] != "T") throw Error(DEMON_ERR_ARG, "not a FITS file");
    return out;
}

const Hdu &find_table(const std::vector<Hdu> &hdus, const std::string &name) {
    for (const Hdu &h : hdus) {
        auto it = h.kv.find("EXTNAME");
        if (it != h.kv.end() && it->second == name) return h;
    }
    throw Error(DEMON_ERR_ARG, "FITS file has no " + name + " table");
}

// Byte offset and element count of each column (TFORM = rT, T = D only here).
void columns(const Hdu &h, std::vector<long> &off, std::vector<long> &cnt) {
    const long long nf = as_int(h, "TFIELDS");
    long o = 0;
    for (long long k = 1; k <= nf; ++k) {
        const std::string tf = h.kv.at("TFORM" + std::to_string(k));
        size_t p = 0;
        while (p < tf.size() && std::isdigit((unsigned char)tf[p])) ++p;
        const long r = p ? std::stol(tf.substr(0, p)) : 1;
        if (p >= tf.size() || tf[p] != 'D') throw Error(DEMON_ERR_ARG, "FITS column type " + tf + " is not D");
        off.push_back(o);
        cnt.push_back(r);
        o += 8 * r;
    }
}

}  // namespace

double FitsCloud::key(const std::string &name, double fallback) const {
    for (const FitsKey &k : keys)
        if (upper(k.name) == upper(name)) return k.value;
    return fallback;
}

FitsWriter::FitsWriter(const std::string &path, long flags, const std::vector<FitsKey> &keys, int64_t n,
                       const double *mass, const double *charge, const double *x, const double *y,
                       const double *vx, const double *vy)
    : n_(n) {
    if (n < 1) throw Error(DEMON_ERR_ARG, "FITS: no particles");
    f_ = std::fopen(path.c_str(), "w+b");
    if (!f_) throw Error(DEMON_ERR_ARG, "cannot create " + path);
    // primary HDU: empty 16-bit image, FORCES, the forces' keys
    std::vector<std::string> c;
    c.push_back(value_card("SIMPLE", "T", "file does conform to FITS standard"));
    c.push_back(int_card("BITPIX", 16, "number of bits per data pixel"));
    c.push_back(int_card("NAXIS", 0, "number of data axes"));
    c.push_back(value_card("EXTEND", "T", "FITS dataset may contain extensions"));
    c.push_back(int_card("FORCES", flags, "Force configuration."));
    for (const FitsKey &k : keys) c.push_back(dbl_card(k.name, k.value, k.comment));
    const std::string prim = header_bytes(c);
    write_all(f_, prim.data(), prim.size(), "primary header");
    // CLOUD: MASS, CHARGE
    const std::string cloud = header_bytes(table_header(16, n, {"MASS", "CHARGE"}, {"D", "D"}, {"kg", "C"}, "CLOUD"));
    write_all(f_, cloud.data(), cloud.size(), "CLOUD header");
    std::vector<unsigned char> d((size_t)padded(16 * (long)n), 0);
    for (int64_t i = 0; i < n; ++i) {
        put_be(&d[(size_t)(16 * i)], mass[i]);
        put_be(&d[(size_t)(16 * i + 8)], charge[i]);
    }
    write_all(f_, d.data(), d.size(), "CLOUD data");
    // TIME_STEP: TIME, 4 vectors of n
    const std::string vec = std::to_string(n) + "D";
    const long width = 8 + 32 * (long)n;
    const long ts_header = std::ftell(f_);
    const std::string ts = header_bytes(table_header(width, 0, {"TIME", "X_POSITION", "Y_POSITION", "X_VELOCITY", "Y_VELOCITY"},
                                                     {"D", vec, vec, vec, vec}, {"s", "m", "m", "m/s", "m/s"}, "TIME_STEP"));
    write_all(f_, ts.data(), ts.size(), "TIME_STEP header");
    naxis2_card_ = ts_header + 4 * kCard;  // XTENSION, BITPIX, NAXIS, NAXIS1, then NAXIS2
    data_start_ = ts_header + (long)ts.size();
    std::fflush(f_);
    writeTimeStep(0.0, x, y, vx, vy);  // Cloud.cpp:234-241: the first row at time 0.0
}

FitsWriter::FitsWriter(const std::string &path) {
    f_ = std::fopen(path.c_str(), "r+b");
    if (!f_) throw Error(DEMON_ERR_ARG, "cannot open " + path);
    const std::vector<Hdu> hdus = read_all(f_);
    const Hdu &ts = find_table(hdus, "TIME_STEP");
    if (&ts != &hdus.back()) throw Error(DEMON_ERR_ARG, "TIME_STEP is not the last HDU of " + path);
    std::vector<long> off, cnt;
    columns(ts, off, cnt);
    if (cnt.size() != 5 || cnt[0] != 1 || as_int(ts, "NAXIS1") != 8 + 32 * cnt[1])
        throw Error(DEMON_ERR_ARG, "unexpected TIME_STEP layout in " + path);
    n_ = cnt[1];
    rows_ = as_int(ts, "NAXIS2");
    naxis2_card_ = ts.card_at.at("NAXIS2");
    data_start_ = ts.data_start;
}

FitsWriter::~FitsWriter() {
    if (f_) std::fclose(f_);
}

void FitsWriter::update_rows() {
    const std::string c = int_card("NAXIS2", rows_, "number of rows in table");
    std::fseek(f_, naxis2_card_, SEEK_SET);
    write_all(f_, c.data(), c.size(), "NAXIS2");
}

void FitsWriter::writeTimeStep(double t, const double *x, const double *y, const double *vx, const double *vy) {
    const long width = 8 + 32 * (long)n_;
    std::vector<unsigned char> row((size_t)width);
    put_be(row.data(), t);
    const double *src[4] = {x, y, vx, vy};
    for (int a = 0; a < 4; ++a)
        for (int64_t i = 0; i < n_; ++i) put_be(&row[(size_t)(8 + 8 * (a * n_ + i))], src[a][i]);
    const long at = data_start_ + rows_ * width;
    std::fseek(f_, at, SEEK_SET);
    write_all(f_, row.data(), row.size(), "TIME_STEP row");
    ++rows_;
    // zero padding to the block boundary, then the row count
    const long end = data_start_ + rows_ * width;
    const std::vector<unsigned char> zeros((size_t)(padded(end) - end), 0);
    write_all(f_, zeros.data(), zeros.size(), "padding");
    update_rows();
    std::fflush(f_);
}

FitsCloud readFits(const std::string &path, int64_t row) {
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(DEMON_ERR_ARG, "cannot open " + path);
    FitsCloud out;
    try {
        const std::vector<Hdu> hdus = read_all(f);
        const Hdu &prim = hdus[0];
        if (prim.kv.count("FORCES")) out.forces = (long)as_int(prim, "FORCES");
        for (const std::string &k : prim.order) {
            if (k == "SIMPLE" || k == "BITPIX" || k == "NAXIS" || k == "EXTEND" || k == "FORCES") continue;
            char *end = nullptr;
            const std::string &v = prim.kv.at(k);
            const double d = std::strtod(v.c_str(), &end);
            if (end && *end == '\0' && end != v.c_str()) out.keys.push_back({k, d, prim.comment.at(k)});
        }
        const Hdu &cl = find_table(hdus, "CLOUD");
        std::vector<long> off, cnt;
        columns(cl, off, cnt);
        out.n = as_int(cl, "NAXIS2");
        const long cw = (long)as_int(cl, "NAXIS1");
        std::vector<unsigned char> d((size_t)(cw * out.n));
        std::fseek(f, cl.data_start, SEEK_SET);
        if (std::fread(d.data(), 1, d.size(), f) != d.size()) throw Error(DEMON_ERR_ARG, "truncated CLOUD table");
        out.mass.resize((size_t)out.n);
        out.charge.resize((size_t)out.n);
        for (int64_t i = 0; i < out.n; ++i) {
            out.mass[(size_t)i] = get_be(&d[(size_t)(cw * i + off[0])]);
            out.charge[(size_t)i] = get_be(&d[(size_t)(cw * i + off[1])]);
        }
        const Hdu &ts = find_table(hdus, "TIME_STEP");
        std::vector<long> toff, tcnt;
        columns(ts, toff, tcnt);
        if (tcnt.size() != 5 || tcnt[1] != out.n) throw Error(DEMON_ERR_ARG, "TIME_STEP does not match CLOUD");
        out.rows = as_int(ts, "NAXIS2");
        if (out.rows < 1) throw Error(DEMON_ERR_ARG, "TIME_STEP has no rows");
        const int64_t r = row < 0 ? out.rows - 1 : row;
        if (r >= out.rows) throw Error(DEMON_ERR_ARG, "TIME_STEP row out of range");
        const long tw = (long)as_int(ts, "NAXIS1");
        std::vector<unsigned char> rb((size_t)tw);
        std::fseek(f, ts.data_start + r * tw, SEEK_SET);
        if (std::fread(rb.data(), 1, rb.size(), f) != rb.size()) throw Error(DEMON_ERR_ARG, "truncated TIME_STEP row");
        out.time = get_be(&rb[(size_t)toff[0]]);
        std::vector<double> *dst[4] = {&out.x, &out.y, &out.vx, &out.vy};
        for (int a = 0; a < 4; ++a) {
            dst[a]->resize((size_t)out.n);
            for (int64_t i = 0; i < out.n; ++i) (*dst[a])[(size_t)i] = get_be(&rb[(size_t)(toff[a + 1] + 8 * i)]);
        }
    } catch (...) {
        std::fclose(f);
        throw;
    }
    std::fclose(f);
    return out;
}

FitsSeries readFitsSeries(const std::string &path, int column, int64_t nuse) {
    if (column < 0 || column > 3) throw Error(DEMON_ERR_ARG, "column must be 0..3");
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(DEMON_ERR_ARG, "cannot open " + path);
    FitsSeries out;
    try {
        const std::vector<Hdu> hdus = read_all(f);
        const Hdu &ts = find_table(hdus, "TIME_STEP");
        std::vector<long> off, cnt;
        columns(ts, off, cnt);
        if (cnt.size() != 5) throw Error(DEMON_ERR_ARG, "TIME_STEP is not a DEMON table");
        out.n = cnt[1];
        out.rows = as_int(ts, "NAXIS2");
        out.nuse = (nuse < 1 || nuse > out.n) ? out.n : nuse;
        const long tw = (long)as_int(ts, "NAXIS1");
        std::vector<unsigned char> rb((size_t)tw);
        out.data.resize((size_t)(out.rows * out.nuse));
        std::fseek(f, ts.data_start, SEEK_SET);
        for (int64_t r = 0; r < out.rows; ++r) {
            if (std::fread(rb.data(), 1, rb.size(), f) != rb.size()) throw Error(DEMON_ERR_ARG, "truncated TIME_STEP row");
            for (int64_t i = 0; i < out.nuse; ++i)
                out.data[(size_t)(r * out.nuse + i)] = get_be(&rb[(size_t)(off[column + 1] + 8 * i)]);
        }
    } catch (...) {
        std::fclose(f);
        throw;
    }
    std::fclose(f);
    return out;
}

}  // namespace demon_amd
