// Content identities of the dense stage's inputs (host code, no device work).
//
// request_content_hash (crates/video-to-3d-core/src/reconstruction_stages.rs:458-510) is the
// identity a DenseStageArtifact records as `request_hash`: SHA-256 over a domain tag, every
// ReconstructionOptions field in little-endian form, and each frame's size and RGBA bytes,
// printed as "sha256:" and 64 lowercase hex digits (ContentHash::of_digest,
// scene_project.rs:473-475). ContentHash::of_bytes (scene_project.rs:463-465) is the same
// digest over a caller's bytes: the artifact's `cameras_hash` is that of the camera stage's
// JSON document.
//
// SHA-256 itself follows FIPS 180-4 (sections 4.1.2, 4.2.2, 5.1.1, 5.3.3 and 6.2).

#include <cstdint>
#include <cstring>

#include "../../include/vt3d_dense.h"

namespace {

// FIPS 180-4 4.2.2: the first 32 bits of the fractional parts of the cube roots of the first
// 64 primes.
constexpr uint32_t kRound[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4,
    0xab1c5ed5, 0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe,
    0x9bdc06a7, 0xc19bf174, 0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f,
    0x4a7484aa, 0x5cb0a9dc, 0x76f988da, 0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7,
    0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967, 0x27b70a85, 0x2e1b2138, 0x4d2c6dfc,
    0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85, 0xa2bfe8a1, 0xa81a664b,
    0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070, 0x19a4c116,
    0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7,
    0xc67178f2};

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

struct Sha256 {
    // FIPS 180-4 5.3.3: the first 32 bits of the fractional parts of the square roots of the
    // first 8 primes.
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                     0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    uint8_t block[64];
    size_t fill = 0;
    uint64_t length = 0;  // bytes hashed

    void compress(const uint8_t *p) {
        uint32_t w[64];
        for (int t = 0; t < 16; ++t)
            w[t] = (uint32_t)p[4 * t] << 24 | (uint32_t)p[4 * t + 1] << 16 |
                   (uint32_t)p[4 * t + 2] << 8 | (uint32_t)p[4 * t + 3];
        for (int t = 16; t < 64; ++t) {
            const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
            const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
        for (int t = 0; t < 64; ++t) {
            const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
            const uint32_t ch = (e & f) ^ (~e & g);
            const uint32_t t1 = k + S1 + ch + kRound[t] + w[t];
            const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
            const uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
            const uint32_t t2 = S0 + maj;
            k = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = b;
            b = a;
            a = t1 + t2;
        }
        h[0] += a;
        h[1] += b;
        h[2] += c;
        h[3] += d;
        h[4] += e;
        h[5] += f;
        h[6] += g;
        h[7] += k;
    }

    void update(const void *data, size_t n) {
        const uint8_t *p = static_cast<const uint8_t *>(data);
        length += n;
        if (fill) {
            const size_t take = n < 64 - fill ? n : 64 - fill;
            std::memcpy(block + fill, p, take);
            fill += take;
            p += take;
            n -= take;
            if (fill < 64) return;
            compress(block);
            fill = 0;
        }
        for (; n >= 64; n -= 64, p += 64) compress(p);
        std::memcpy(block, p, n);
        fill = n;
    }

    template <class T> void update_le(T v) {
        uint8_t b[sizeof(T)];
        for (size_t i = 0; i < sizeof(T); ++i) b[i] = (uint8_t)((uint64_t)v >> (8 * i));
        update(b, sizeof(T));
    }

    // FIPS 180-4 5.1.1 padding, then the digest as "sha256:" + 64 lowercase hex digits.
    void finish(char *out) {
        const uint64_t bits = length * 8;
        const uint8_t one = 0x80, zero = 0;
        update(&one, 1);
        while (fill != 56) update(&zero, 1);
        uint8_t len[8];
        for (int i = 0; i < 8; ++i) len[i] = (uint8_t)(bits >> (56 - 8 * i));
        update(len, 8);
        static const char digits[] = "0123456789abcdef";
        std::memcpy(out, "sha256:", 7);
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 4; ++j) {
                const uint8_t byte = (uint8_t)(h[i] >> (24 - 8 * j));
                out[7 + 8 * i + 2 * j] = digits[byte >> 4];
                out[7 + 8 * i + 2 * j + 1] = digits[byte & 15];
            }
        out[71] = '\0';
    }
};

uint32_t f32_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// REQUEST_HASH_DOMAIN, reconstruction_stages.rs:40 (with its terminating NUL).
constexpr char kRequestDomain[] = "video-to-3d/reconstruction-request/v1";

}  // namespace

extern "C" {

void vt3d_reconstruction_options_default(vt3d_reconstruction_options *o) {
    // ReconstructionOptions::default, reconstruction/types.rs:37-50.
    if (!o) return;
    *o = vt3d_reconstruction_options{};
    o->max_features = 320;
    o->min_feature_distance = 7;
    o->descriptor_radius = 3;
    o->match_radius = 42;
    o->max_descriptor_distance = 36.0f;
    o->ratio_threshold = 0.82f;
    o->has_focal_length_pixels = 0;
    o->max_dense_working_set_bytes = 256ull * 1024 * 1024;  // types.rs:10
}

void vt3d_content_hash(const uint8_t *bytes, size_t n, char *out) {
    Sha256 s;
    if (n) s.update(bytes, n);
    s.finish(out);
}

int32_t vt3d_request_content_hash(const vt3d_frame *frames, size_t n_frames,
                                  const vt3d_reconstruction_options *options, vt3d_poll_fn poll,
                                  void *poll_user, char *out) {
    if (!out || !options || (n_frames && !frames)) return VT3D_E_INVALID;
    for (size_t i = 0; i < n_frames; ++i)
        if (!frames[i].rgba && frames[i].width && frames[i].height) return VT3D_E_INVALID;
    const vt3d_reconstruction_options &o = *options;
    Sha256 s;
    s.update(kRequestDomain, sizeof(kRequestDomain));  // the tag and its NUL
    s.update_le<uint64_t>(o.max_features);
    s.update_le<uint32_t>(o.min_feature_distance);
    s.update_le<uint32_t>(o.descriptor_radius);
    s.update_le<uint32_t>(o.match_radius);
    s.update_le<uint32_t>(f32_bits(o.max_descriptor_distance));
    s.update_le<uint32_t>(f32_bits(o.ratio_threshold));
    if (o.has_focal_length_pixels) {
        s.update_le<uint8_t>(1);
        s.update_le<uint32_t>(f32_bits(o.focal_length_pixels));
    } else {
        s.update_le<uint8_t>(0);
    }
    s.update_le<uint64_t>(o.max_dense_working_set_bytes);
    s.update_le<uint64_t>(n_frames);
    bool canceled = false;  // ReconstructionControl: sticky
    for (size_t i = 0; i < n_frames; ++i) {
        const vt3d_frame &f = frames[i];
        const uint64_t len = 4ull * f.width * f.height;
        s.update_le<uint32_t>(f.width);
        s.update_le<uint32_t>(f.height);
        s.update_le<uint64_t>(len);
        // the pixels in 64 KiB chunks, a cancellation check after each
        for (uint64_t at = 0; at < len; at += 64 * 1024) {
            const uint64_t n = len - at < 64 * 1024 ? len - at : 64 * 1024;
            s.update(f.rgba + at, (size_t)n);
            canceled = canceled || (poll && poll(poll_user) != 0);
            if (canceled) return VT3D_E_CANCELED;
        }
    }
    s.finish(out);
    return VT3D_OK;
}

}  // extern "C"
