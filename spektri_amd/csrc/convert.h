// Input formats without a loader of their own in the transform kernels
// (spektri/src/inputformats.rs:65-155): one elementwise pass turns the raw
// stream into the form a loader reads -- s16 (real s8, s16be), f32 (real u8,
// f32be) or cf32 (cs8, cu8, cs16be, cf32be) -- with the reference's
// per-sample arithmetic: integers convert exactly, u8/cu8 subtract 127.4 in
// f32 (inputformats.rs:70,108-109), big-endian formats swap bytes.  s16le,
// f32le, cs16le and cf32le are read in place and never pass through here.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spk {

// Conversion kinds (host and device agree on these numbers).
// A complex sample is two numbers, so each kind serves the real and the
// complex format of the same number type.
enum ConvKind : int {
  kConvS8ToS16 = 1,     // s8: 4 numbers per raw word -> 4 s16
  kConvS16beToS16 = 2,  // s16be: 2 numbers -> 2 s16
  kConvU8ToF32 = 3,     // u8, cu8: 4 numbers -> 4 f32 (x - 127.4)
  kConvS8ToF32 = 4,     // cs8: 4 numbers -> 4 f32
  kConvF32beToF32 = 5,  // f32be, cf32be: 1 number -> 1 f32
  kConvS16beToF32 = 7,  // cs16be: 2 numbers -> 2 f32
};

struct ConvParams {
  const uint32_t* src;  // raw stream, 4-byte aligned
  void* dst;            // converted stream (device buffer)
  long long words;      // whole raw 32-bit words to convert
  int tail;             // raw bytes after them (1-byte kinds only, 0..3), read one by one
  int kind;
};

__device__ __forceinline__ float u8_sample(uint32_t b) { return (float)b - 127.4f; }
__device__ __forceinline__ float s8_sample(uint32_t b) { return (float)(int8_t)(uint8_t)b; }
__device__ __forceinline__ uint32_t bswap16x2(uint32_t v) {
  return ((v & 0x00FF00FFu) << 8) | ((v >> 8) & 0x00FF00FFu);
}

// One raw word per thread per iteration (a wave reads 256 contiguous bytes);
// the output of word i starts at byte i * (output bytes per raw word).  The
// 1-3 bytes of a stream of 1-byte numbers that end inside a word are read
// byte by byte (never past the stream) by the first thread.
__global__ void convert_kernel(ConvParams P) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  if (P.tail > 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    const uint8_t* b = reinterpret_cast<const uint8_t*>(P.src + P.words);
    for (int j = 0; j < P.tail; ++j) {
      const long long o = 4 * P.words + j;
      const uint32_t v = b[j];
      if (P.kind == kConvS8ToS16) static_cast<int16_t*>(P.dst)[o] = (int16_t)(int8_t)(uint8_t)v;
      else if (P.kind == kConvU8ToF32) static_cast<float*>(P.dst)[o] = u8_sample(v);
      else if (P.kind == kConvS8ToF32) static_cast<float*>(P.dst)[o] = s8_sample(v);
    }
  }
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P.words; i += stride) {
    const uint32_t v = P.src[i];
    switch (P.kind) {
      case kConvS8ToS16: {
        const uint32_t a = (uint32_t)(uint16_t)(int16_t)(int8_t)(v & 0xFF);
        const uint32_t b = (uint32_t)(uint16_t)(int16_t)(int8_t)((v >> 8) & 0xFF);
        const uint32_t c = (uint32_t)(uint16_t)(int16_t)(int8_t)((v >> 16) & 0xFF);
        const uint32_t d = (uint32_t)(uint16_t)(int16_t)(int8_t)(v >> 24);
        static_cast<uint2*>(P.dst)[i] = make_uint2(a | (b << 16), c | (d << 16));
        break;
      }
      case kConvS16beToS16:
        static_cast<uint32_t*>(P.dst)[i] = bswap16x2(v);
        break;
      case kConvU8ToF32:
        static_cast<float4*>(P.dst)[i] =
            make_float4(u8_sample(v & 0xFF), u8_sample((v >> 8) & 0xFF), u8_sample((v >> 16) & 0xFF), u8_sample(v >> 24));
        break;
      case kConvS8ToF32:
        static_cast<float4*>(P.dst)[i] =
            make_float4(s8_sample(v & 0xFF), s8_sample((v >> 8) & 0xFF), s8_sample((v >> 16) & 0xFF), s8_sample(v >> 24));
        break;
      case kConvF32beToF32:
        static_cast<uint32_t*>(P.dst)[i] = __builtin_bswap32(v);
        break;
      case kConvS16beToF32: {
        const uint32_t s = bswap16x2(v);
        static_cast<float2*>(P.dst)[i] = make_float2((float)(int16_t)(s & 0xFFFF), (float)(int16_t)(s >> 16));
        break;
      }
      default:
        break;
    }
  }
}

}  // namespace spk
