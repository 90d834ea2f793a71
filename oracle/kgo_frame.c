/*
 * kgo_frame.c -- CPU oracle of the exchange frame (the unit of local
 * shuffle files and the wire): its 44-byte header and the CRC32C of its
 * body.
 *
 * TEST INFRASTRUCTURE ONLY (see kgo.h): the checker of libkgpu's frame
 * writer.  The frame's Arrow IPC body is checked with pyarrow's IPC reader
 * (an independent implementation of the Arrow format), not restated here.
 *
 * Restates crates/krishiv-shuffle/src/frame.rs:1-26 (layout), 118-150
 * (header_bytes), 156-201 (decode), 251-274 (crc32c).
 */
#include <string.h>

#include "kgo.h"

/* crc32c (frame.rs:251-274): reflected, polynomial 0x82F63B78, initial and
 * final value all ones, one byte a step through a 256-entry table. */
uint32_t kgo_crc32c(const uint8_t *p, uint64_t n) {
  static uint32_t table[256];
  static int ready = 0;
  if (!ready) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ 0x82F63B78u : c >> 1;
      table[i] = c;
    }
    ready = 1;
  }
  uint32_t crc = 0xFFFFFFFFu;
  for (uint64_t i = 0; i < n; i++) crc = table[(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
  return ~crc;
}

static void put32(uint8_t *o, uint32_t v) { memcpy(o, &v, 4); }
static uint32_t get32(const uint8_t *p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}

/* header_bytes (frame.rs:132-150) */
void kgo_frame_header(const kgo_frame_hdr *h, const uint8_t *body, uint64_t len, uint8_t out[44]) {
  memcpy(out, "KXFR", 4);
  out[4] = 1; /* FORMAT_VERSION, little-endian u16 */
  out[5] = 0;
  out[6] = (uint8_t)h->kind;
  out[7] = 0; /* no compression */
  memcpy(out + 8, &h->id, 8);
  put32(out + 16, h->stage_attempt);
  put32(out + 20, h->map);
  put32(out + 24, h->attempt);
  put32(out + 28, h->sequence);
  put32(out + 32, h->reduce);
  put32(out + 36, (uint32_t)len);
  put32(out + 40, kgo_crc32c(body, len));
}

/* decode (frame.rs:156-201): 0 and the header and body range, or one of
 * KGO_FRAME_E*. */
int kgo_frame_decode(const uint8_t *f, uint64_t n, kgo_frame_hdr *h, uint64_t *body_off,
                     uint64_t *body_len) {
  if (n < 44) return KGO_FRAME_ETRUNCATED;
  if (memcmp(f, "KXFR", 4) != 0) return KGO_FRAME_EMAGIC;
  if (f[4] != 1 || f[5] != 0) return KGO_FRAME_EVERSION;
  if (f[6] < 1 || f[6] > 3) return KGO_FRAME_EKIND;
  if (f[7] != 0) return KGO_FRAME_ECOMPRESSION;
  uint64_t len = get32(f + 36);
  if (n < 44 + len) return KGO_FRAME_ETRUNCATED;
  if (n != 44 + len) return KGO_FRAME_EBODY;
  if (kgo_crc32c(f + 44, len) != get32(f + 40)) return KGO_FRAME_ECHECKSUM;
  h->kind = f[6];
  memcpy(&h->id, f + 8, 8);
  h->stage_attempt = get32(f + 16);
  h->map = get32(f + 20);
  h->attempt = get32(f + 24);
  h->sequence = get32(f + 28);
  h->reduce = get32(f + 32);
  *body_off = 44;
  *body_len = len;
  return 0;
}
