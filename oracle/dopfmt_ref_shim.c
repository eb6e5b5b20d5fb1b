/*
 * Wrapper exposing the reference firmware's DOp template patching (patch_dop in
 * fw/ublaze/src/hpu_dop_fmt.h:206-229) through a flat C function, so tests can
 * compare this build's patching against the reference's own C code.  The header is
 * compiled from the reference checkout in place; nothing from it is copied here.
 * Test infrastructure only.
 */
#include <string.h>
#include "hpu_dop_fmt.h"

/* dst_cid/src_cid: base cid per operand; imm: [nimm][IMMEDIAT_MSG_MAX_SLOT] packed messages */
uint32_t dopref_patch(uint32_t raw, const uint16_t* dst_cid, int ndst, const uint16_t* src_cid, int nsrc,
                      const uint16_t* imm, int nimm) {
  OperandBundle_t dst, src;
  ImmediatBundle_t im;
  memset(&dst, 0, sizeof(dst));
  memset(&src, 0, sizeof(src));
  memset(&im, 0, sizeof(im));
  dst.len = (uint8_t)ndst;
  for (int i = 0; i < ndst && i < OPERAND_BUNDLE_MAX_SLOT; i++) dst.operand[i].cid_ofst = dst_cid[i];
  src.len = (uint8_t)nsrc;
  for (int i = 0; i < nsrc && i < OPERAND_BUNDLE_MAX_SLOT; i++) src.operand[i].cid_ofst = src_cid[i];
  im.len = (uint8_t)nimm;
  for (int i = 0; i < nimm && i < IMMEDIAT_BUNDLE_MAX_SLOT; i++)
    for (int j = 0; j < IMMEDIAT_MSG_MAX_SLOT; j++) im.cst[i].msg[j] = imm[i * IMMEDIAT_MSG_MAX_SLOT + j];
  DOpu_t d;
  d.raw = raw;
  patch_dop(&d, &dst, &src, &im);
  return d.raw;
}

int dopref_heap_start(void) { return HEAP_START_SLOT; }
