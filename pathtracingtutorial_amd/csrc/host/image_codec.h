// 8-bit image codecs for the scene loaders and the image writer.
//
// Decoders (PNG, JPEG baseline and progressive) stand where the reference
// calls stb_image (io/image_io.cpp:47-64 for .png/.jpg files,
// io/gltf_loader.cpp:22-40 for glTF images); encoders (PNG, JPEG, BMP, TGA)
// where it calls stb_image_write (io/image_io.cpp:129-141).  Both are written
// from the format specifications (PNG: RFC 2083 / RFC 1950-1951; JPEG:
// ITU T.81), not from stb.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace ptamd {

// RGBA, 8 bits per channel, row 0 = the image's top row.
struct Image8 {
    int width = 0, height = 0;
    std::vector<uint8_t> rgba;
};

bool read_file(const std::string& path, std::vector<uint8_t>& out, std::string* err);

// Decodes PNG or JPEG (chosen by the signature) to RGBA8 (grey and RGB get
// alpha 255).  False with *err on malformed or unsupported data.
bool decode_image(const uint8_t* data, size_t n, Image8& out, std::string* err);
bool decode_png(const uint8_t* data, size_t n, Image8& out, std::string* err);
bool decode_jpeg(const uint8_t* data, size_t n, Image8& out, std::string* err);

// Encoders of 8-bit RGB (w*h*3 bytes, row 0 = top).
void encode_png(int w, int h, const uint8_t* rgb, std::vector<uint8_t>& out);
void encode_jpeg(int w, int h, const uint8_t* rgb, int quality, std::vector<uint8_t>& out);
void encode_bmp(int w, int h, const uint8_t* rgb, std::vector<uint8_t>& out);
void encode_tga(int w, int h, const uint8_t* rgb, std::vector<uint8_t>& out);

}  // namespace ptamd
