// KITTI stereo loader and PNG reader behind include/vxh/datasets.h.
//
//   directory listing  vx-vision/src/datasets/kitti.rs:26-41: image_2 when it
//                      exists, else image_0; every entry whose extension is
//                      exactly "png", sorted byte-wise by name.
//   poses              kitti.rs:62-87: one pose per line that holds at least 12
//                      numbers after dropping tokens that do not parse as f32;
//                      rotation = values 0 1 2 4 5 6 8 9 10, translation = 3 7 11;
//                      frame i takes the i-th accepted line.
//   PNG                the PNG specification's chunk layout, zlib stream and
//                      five row filters (none, sub, up, average, Paeth), for
//                      8-bit gray, gray + alpha, RGB and RGBA without interlace.
#include <dirent.h>
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <new>
#include <string>
#include <vector>

#include "sgm_pipeline.h"
#include "vxh/datasets.h"

namespace {

vxh_status invalid(const std::string &msg) {
    return vxh::report_error(VXH_ERR_INVALID_CONFIG, "invalid config: " + msg);
}

bool exists(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

bool is_file(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

bool read_file(const std::string &path, std::string *out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out->assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return !f.bad();
}

// Rust's str::lines: split on '\n', drop one trailing '\r' per line, no
// final empty line.
std::vector<std::string> lines_of(const std::string &text) {
    std::vector<std::string> out;
    size_t start = 0;
    while (start < text.size()) {
        size_t end = text.find('\n', start);
        if (end == std::string::npos) end = text.size();
        std::string line = text.substr(start, end - start);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        out.push_back(line);
        start = end + 1;
    }
    return out;
}

std::vector<std::string> split_ws(const std::string &line) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
        size_t j = i;
        while (j < line.size() && !std::isspace((unsigned char)line[j])) ++j;
        if (j > i) out.push_back(line.substr(i, j - i));
        i = j;
    }
    return out;
}

// A whole token as f32, as Rust's str::parse::<f32> accepts it: decimal
// (with optional sign, fraction and exponent) or inf / infinity / nan in any
// case; hexadecimal and trailing characters are rejected.
bool parse_f32(const std::string &tok, float *v) {
    if (tok.empty()) return false;
    for (char ch : tok)
        if (ch == 'x' || ch == 'X' || ch == '(' || ch == ')') return false;
    char *end = nullptr;
    errno = 0;
    const float f = std::strtof(tok.c_str(), &end);
    if (end != tok.c_str() + tok.size()) return false;
    *v = f;   // overflow gives +-inf, as Rust's parse does
    return true;
}

// ---- PNG ----

// Largest decoded image accepted (a header cannot make the reader allocate
// more than this).
constexpr uint64_t kMaxImageBytes = 1ull << 30;

struct Image {
    uint32_t w = 0, h = 0, channels = 0;
    std::vector<uint8_t> pix;
};

uint32_t be32(const uint8_t *p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

vxh_status png_load(const std::string &path, Image *img, bool header_only) {
    std::string file;
    if (!read_file(path, &file)) return invalid("cannot read " + path);
    const uint8_t *d = (const uint8_t *)file.data();
    const size_t n = file.size();
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n < 8 || std::memcmp(d, sig, 8) != 0) return invalid(path + ": not a PNG file");
    size_t pos = 8;
    bool have_header = false;
    std::vector<uint8_t> zdata;
    while (true) {
        if (pos + 12 > n) return invalid(path + ": truncated PNG");
        const uint32_t len = be32(d + pos);
        if (len > n - pos - 12) return invalid(path + ": truncated PNG chunk");
        const uint8_t *type = d + pos + 4, *data = d + pos + 8;
        const uint32_t crc = be32(data + len);
        if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4) != crc)
            return invalid(path + ": PNG chunk CRC mismatch");
        const std::string t((const char *)type, 4);
        if (!have_header) {
            if (t != "IHDR" || len != 13) return invalid(path + ": PNG without IHDR");
            img->w = be32(data);
            img->h = be32(data + 4);
            const uint8_t depth = data[8], color = data[9];
            if (img->w == 0 || img->h == 0 || img->w > 65535 || img->h > 65535)
                return invalid(path + ": PNG size out of range");
            if (data[10] != 0 || data[11] != 0) return invalid(path + ": unknown PNG method");
            if (data[12] != 0) return invalid(path + ": interlaced PNG not supported");
            if (depth != 8) return invalid(path + ": only 8-bit PNG samples are supported");
            switch (color) {
                case 0: img->channels = 1; break;
                case 4: img->channels = 2; break;
                case 2: img->channels = 3; break;
                case 6: img->channels = 4; break;
                default: return invalid(path + ": palette PNG not supported");
            }
            if ((uint64_t)img->w * img->h * img->channels > kMaxImageBytes)
                return invalid(path + ": PNG image over 1 GiB");
            have_header = true;
            if (header_only) return VXH_OK;
        } else if (t == "IDAT") {
            zdata.insert(zdata.end(), data, data + len);
        } else if (t == "IEND") {
            break;
        }
        pos += 12 + (size_t)len;
    }
    if (!have_header) return invalid(path + ": PNG without IHDR");
    const size_t row = (size_t)img->w * img->channels;
    const size_t expect = (row + 1) * img->h;
    std::vector<uint8_t> raw;
    try {
        raw.resize(expect);
        img->pix.assign(row * img->h, 0);
    } catch (const std::bad_alloc &) {
        return vxh::report_error(VXH_ERR_BUFFER_ALLOC, "buffer allocation failed (PNG " + path + ")");
    }
    uLongf got = (uLongf)expect;
    if (uncompress(raw.data(), &got, zdata.data(), (uLong)zdata.size()) != Z_OK || got != expect)
        return invalid(path + ": corrupt PNG image data");
    const int bpp = (int)img->channels;
    for (uint32_t y = 0; y < img->h; ++y) {
        const uint8_t f = raw[y * (row + 1)];
        const uint8_t *src = &raw[y * (row + 1) + 1];
        uint8_t *cur = &img->pix[y * row];
        const uint8_t *prev = y ? &img->pix[(y - 1) * row] : nullptr;
        for (size_t i = 0; i < row; ++i) {
            const int a = i >= (size_t)bpp ? cur[i - bpp] : 0;
            const int b = prev ? prev[i] : 0;
            const int c = (prev && i >= (size_t)bpp) ? prev[i - bpp] : 0;
            int pred;
            switch (f) {
                case 0: pred = 0; break;
                case 1: pred = a; break;
                case 2: pred = b; break;
                case 3: pred = (a + b) >> 1; break;
                case 4: pred = paeth(a, b, c); break;
                default: return invalid(path + ": unknown PNG row filter");
            }
            cur[i] = (uint8_t)(src[i] + pred);
        }
    }
    return VXH_OK;
}

// DynamicImage::to_luma8 of the `image` crate 0.25: Rec. 709 weights in
// integers, truncating division; alpha is dropped.
void to_luma8(const Image &img, uint8_t *dst) {
    const size_t n = (size_t)img.w * img.h;
    const uint8_t *p = img.pix.data();
    const uint32_t c = img.channels;
    for (size_t i = 0; i < n; ++i, p += c)
        dst[i] = c <= 2 ? p[0]
                        : (uint8_t)((2126u * p[0] + 7152u * p[1] + 722u * p[2]) / 10000u);
}

bool parse_projection(const std::vector<std::string> &lines, const char *label, double P[12]) {
    const std::string want = std::string(label) + ":";
    for (const std::string &line : lines) {
        const std::vector<std::string> tok = split_ws(line);
        if (tok.size() < 13 || tok[0] != want) continue;
        for (int i = 0; i < 12; ++i) {
            char *end = nullptr;
            P[i] = std::strtod(tok[1 + i].c_str(), &end);
            if (end != tok[1 + i].c_str() + tok[1 + i].size()) return false;
        }
        return true;
    }
    return false;
}

}  // namespace

struct vxh_kitti_stereo {
    bool color = false;
    std::vector<std::string> left, right;
    std::vector<std::array<float, 12>> poses;
    vxh_stereo_calib calib{};
};

extern "C" {

vxh_status vxh_kitti_stereo_load(const char *sequence_dir, const char *poses_file,
                                 vxh_kitti_stereo **out) {
    if (!sequence_dir || !out) return invalid("null argument");
    *out = nullptr;
    const std::string dir(sequence_dir);
    auto *ds = new vxh_kitti_stereo();
    ds->color = exists(dir + "/image_2");
    const std::string ldir = dir + (ds->color ? "/image_2" : "/image_0");
    const std::string rdir = dir + (ds->color ? "/image_3" : "/image_1");
    DIR *dh = opendir(ldir.c_str());
    if (!dh) {
        const std::string why = std::strerror(errno);
        delete ds;
        return invalid("cannot read " + ldir + ": " + why);
    }
    std::vector<std::string> names;
    while (dirent *e = readdir(dh)) {
        const std::string name = e->d_name;
        // Path::extension: the part after the last '.', none for ".png" itself.
        const size_t dot = name.rfind('.');
        if (dot != std::string::npos && dot > 0 && name.compare(dot + 1, std::string::npos, "png") == 0)
            names.push_back(name);
    }
    closedir(dh);
    std::sort(names.begin(), names.end());
    for (const std::string &name : names) {
        const std::string r = rdir + "/" + name;
        if (!is_file(r)) {
            delete ds;
            return invalid("missing right image " + r);
        }
        ds->left.push_back(ldir + "/" + name);
        ds->right.push_back(r);
    }
    std::string text;
    if (poses_file && read_file(poses_file, &text)) {
        for (const std::string &line : lines_of(text)) {
            std::vector<float> vals;
            float v;
            for (const std::string &tok : split_ws(line))
                if (parse_f32(tok, &v)) vals.push_back(v);
            if (vals.size() >= 12) {
                std::array<float, 12> a;
                std::copy(vals.begin(), vals.begin() + 12, a.begin());
                ds->poses.push_back(a);
            }
        }
    }
    if (read_file(dir + "/calib.txt", &text)) {
        const std::vector<std::string> lines = lines_of(text);
        double PL[12], PR[12];
        if (parse_projection(lines, ds->color ? "P2" : "P0", PL) &&
            parse_projection(lines, ds->color ? "P3" : "P1", PR) && PL[0] != 0.0) {
            ds->calib.fx = PL[0];
            ds->calib.fy = PL[5];
            ds->calib.cx = PL[2];
            ds->calib.cy = PL[6];
            ds->calib.baseline = (PL[3] - PR[3]) / PL[0];
            ds->calib.valid = 1;
        }
    }
    *out = ds;
    return VXH_OK;
}

void vxh_kitti_stereo_destroy(vxh_kitti_stereo *ds) { delete ds; }

size_t vxh_kitti_stereo_len(const vxh_kitti_stereo *ds) { return ds ? ds->left.size() : 0; }

int32_t vxh_kitti_stereo_is_color(const vxh_kitti_stereo *ds) { return ds && ds->color ? 1 : 0; }

vxh_status vxh_kitti_stereo_frame(const vxh_kitti_stereo *ds, size_t i, vxh_stereo_frame *out) {
    if (!ds || !out) return invalid("null argument");
    if (i >= ds->left.size()) return invalid("frame index out of range");
    std::memset(out, 0, sizeof *out);
    out->timestamp = (double)i / 10.0;
    out->left_path = ds->left[i].c_str();
    out->right_path = ds->right[i].c_str();
    if (i < ds->poses.size()) {
        const std::array<float, 12> &v = ds->poses[i];
        static const int rot[9] = {0, 1, 2, 4, 5, 6, 8, 9, 10};
        for (int k = 0; k < 9; ++k) out->rotation[k] = v[rot[k]];
        out->translation[0] = v[3];
        out->translation[1] = v[7];
        out->translation[2] = v[11];
        out->has_pose = 1;
    }
    return VXH_OK;
}

vxh_status vxh_kitti_stereo_calib(const vxh_kitti_stereo *ds, vxh_stereo_calib *out) {
    if (!ds || !out) return invalid("null argument");
    *out = ds->calib;
    return VXH_OK;
}

vxh_status vxh_png_info(const char *path, uint32_t *w, uint32_t *h, uint32_t *channels) {
    if (!path || !w || !h || !channels) return invalid("null argument");
    Image img;
    vxh_status v = png_load(path, &img, true);
    if (v != VXH_OK) return v;
    *w = img.w;
    *h = img.h;
    *channels = img.channels;
    return VXH_OK;
}

vxh_status vxh_png_read(const char *path, uint8_t *dst, size_t len) {
    if (!path || !dst) return invalid("null argument");
    Image img;
    vxh_status v = png_load(path, &img, false);
    if (v != VXH_OK) return v;
    if (len != img.pix.size()) return invalid("PNG read length");
    std::memcpy(dst, img.pix.data(), len);
    return VXH_OK;
}

vxh_status vxh_png_read_luma8(const char *path, uint8_t *dst, size_t len) {
    if (!path || !dst) return invalid("null argument");
    Image img;
    vxh_status v = png_load(path, &img, false);
    if (v != VXH_OK) return v;
    if (len != (size_t)img.w * img.h) return invalid("PNG read length");
    to_luma8(img, dst);
    return VXH_OK;
}

vxh_status vxh_kitti_stereo_read_pair(const vxh_kitti_stereo *ds, size_t i, uint8_t *left,
                                      uint8_t *right, size_t len) {
    if (!ds || !left || !right) return invalid("null argument");
    if (i >= ds->left.size()) return invalid("frame index out of range");
    vxh_status v = vxh_png_read_luma8(ds->left[i].c_str(), left, len);
    if (v != VXH_OK) return v;
    return vxh_png_read_luma8(ds->right[i].c_str(), right, len);
}

}  // extern "C"
