/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the reference CPU emulator (RPP1011/fast-atari-rs)
 * used as the correctness checker for the HIP frame-step kernels and as the
 * CPU baseline in bench.py.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * Followed files (reference paths):
 *   src/cpu.rs            Cpu, Status, step()            (lines 11-1146)
 *   src/headless_tia.rs   HeadlessTia                    (lines 15-298)
 *   src/pia.rs            Pia (tick_n/read/write)        (lines 25-228)
 *   src/atari.rs          Action, BankScheme,             (lines 9-105)
 *                         HeadlessBus, HeadlessAtari      (lines 384-620)
 *   src/cuda_env.rs       AtariStateGpu::from_headless    (lines 12-94)
 *                         BatchAtariGpu::reset            (lines 131-142)
 *
 * Two deliberate, documented additions (the reference panics or never returns):
 *   - an illegal opcode sets `error = ORC_ERR_ILLEGAL` and ends the frame (the
 *     reference panics at src/cpu.rs:1139-1141; PC has already advanced past
 *     the opcode and the instruction's cycles are not counted, as there);
 *   - the VSYNC-held loop (src/atari.rs:588-590, uncapped in the reference) is
 *     capped at ORC_LOOPA_CAP cycles, after which `error = ORC_ERR_VSYNC_HELD`.
 * An env with a non-zero error is frozen: later frames do nothing.
 *
 * Parity unpinned (stated only by the reference's text, restated here and
 * reviewed, no independent check): the headless TIA's paddle model -- the
 * counter's ramp and the 200 / 50 / 128 paddle values of set_action
 * (src/headless_tia.rs, src/atari.rs:571-582) -- and, on the rendering path
 * (orc_render.inc), the player RESPx formula, the RESMx / RESBL offset and
 * the visible-area resizer constants (src/tia.rs:76-80, 536-555).  Everything
 * else is pinned by tests/test_oracle_*.py (including the flat-array padding
 * reads of src/atari.rs:386-395, 503-507 and the frame-loop caps of
 * src/atari.rs:588-596, by closed forms and the 2600 address map) and
 * tests/golden/.
 *
 * Licence note: the reference is GPL-3.0; this restatement follows its
 * algorithm and is kept under the same licence terms.
 */
#ifndef FA_ORACLE_ORC_H
#define FA_ORACLE_ORC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_ERR_NONE 0
#define ORC_ERR_ILLEGAL 1
#define ORC_ERR_VSYNC_HELD 2
#define ORC_LOOPA_CAP 1000000ull
#define ORC_STATE_BYTES 222

enum { ORC_FIXED = 0, ORC_F8 = 1, ORC_F6 = 2, ORC_F4 = 3 };

/* src/cpu.rs:96-105 */
typedef struct {
    uint16_t pc;
    uint8_t sp, a, x, y, p;
    uint64_t cycles;
} orc_cpu;

/* src/headless_tia.rs:15-88 */
typedef struct {
    uint16_t clock, scanline;
    uint8_t wsync, frame_complete, vsync, vblank;
    uint8_t cxm0p, cxm1p, cxp0fb, cxp1fb, cxm0fb, cxm1fb, cxblpf, cxppmm;
    uint8_t inpt4, inpt5, paddle0, paddle1;
    uint16_t paddle_counter;
    uint8_t paddle_dumped;
    uint16_t pos_p0, pos_p1, pos_m0, pos_m1, pos_bl;
    uint8_t hmp0, hmp1, hmm0, hmm1, hmbl;
    uint8_t grp0, grp1, grp0_old, grp1_old;
    uint8_t enam0, enam1, enabl, enabl_old, vdelp0, vdelp1, vdelbl, resmp0, resmp1;
    uint8_t nusiz0, nusiz1, colup0, colup1, colupf, colubk, ctrlpf, refp0, refp1;
    uint8_t pf0, pf1, pf2;
} orc_tia;

/* src/pia.rs:25-44 */
typedef struct {
    uint8_t ram[128];
    uint8_t pa_out, pa_ddr, pa_in, pb_out, pb_ddr, pb_in;
    uint16_t tval, tps, tsel;
    uint8_t tuf;
} orc_pia;

/* src/atari.rs:398-407 */
typedef struct {
    orc_tia tia;
    orc_pia pia;
    uint8_t mem[8192];
    const uint8_t *rom;
    uint32_t rom_len;
    uint32_t bank;
    int scheme;
} orc_bus;

/* src/atari.rs:556-559 plus the two counters and the error latch above */
typedef struct {
    orc_cpu cpu;
    orc_bus bus;
    int error;
    uint64_t instructions;
    uint64_t last_frame_cycles;
} orc_atari;

/* ---- single-console API ---- */
int orc_scheme_detect(uint32_t rom_len);
void orc_tia_default(orc_tia *t);
void orc_pia_default(orc_pia *p);
void orc_atari_init(orc_atari *e, const uint8_t *rom, uint32_t rom_len);
void orc_atari_set_action(orc_atari *e, uint32_t action);
uint64_t orc_atari_run_frame(orc_atari *e);
void orc_atari_pack(const orc_atari *e, uint8_t *out);
void orc_atari_unpack(orc_atari *e, const uint8_t *in);
uint8_t orc_bus_read(orc_bus *b, uint16_t addr);
void orc_bus_write(orc_bus *b, uint16_t addr, uint8_t val);
int orc_cpu_step_bus(orc_cpu *c, orc_bus *b); /* cycles, or -1 on illegal */
void orc_pia_tick_n(orc_pia *p, uint16_t n);
void orc_pia_tick(orc_pia *p);
uint8_t orc_pia_read(orc_pia *p, uint16_t addr);
void orc_pia_write(orc_pia *p, uint16_t addr, uint8_t val);
void orc_tia_tick_n(orc_tia *t, uint8_t n);
uint16_t orc_tia_skip_to_scanline_end(orc_tia *t);
uint8_t orc_tia_read(const orc_tia *t, uint16_t addr);
void orc_tia_write(orc_tia *t, uint16_t addr, uint8_t val);

/* ---- handle API used from Python (ctypes) ---- */
void *orc_new(const uint8_t *rom, uint32_t rom_len);
void orc_free(void *h);
void orc_reset_frame(void *h); /* BatchAtariGpu::reset for one env: new + run_frame */
uint64_t orc_step(void *h, uint32_t action, uint32_t frames);
void orc_get_state(void *h, uint8_t *out);
void orc_set_state(void *h, const uint8_t *in);
void orc_get_counters(void *h, uint64_t out[4]); /* cpu.cycles, last frame cycles, instructions, error */
void orc_set_cycles(void *h, uint64_t cpu_cycles);
void orc_get_ram(void *h, uint8_t out[128]);
void orc_get_mem(void *h, uint8_t out[8192]);
uint8_t orc_h_bus_read(void *h, uint16_t addr);
void orc_h_bus_write(void *h, uint16_t addr, uint8_t val);

/* ---- batch API (pthreads over contiguous env slices) ---- */
void *orc_batch_new(const uint8_t *rom, uint32_t rom_len, uint32_t n);
void orc_batch_free(void *b);
void orc_batch_reset(void *b, uint32_t nthreads);
/* actions: frames x n, index [f*n + i]; one run_frame per (f, i). */
void orc_batch_run(void *b, const uint8_t *actions, uint32_t frames, uint32_t nthreads);
/* frameskip variant: env i holds action[i] for fs[i] frames (set_action + fs x run_frame). */
void orc_batch_run_fs(void *b, const uint8_t *actions, const uint8_t *fs, uint32_t nthreads);
void orc_batch_states(void *b, uint8_t *out /* n x 222 */);
void orc_batch_set_states(void *b, const uint8_t *in /* n x 222 */);
void orc_batch_counters(void *b, uint64_t *out /* n x 4 */);
void orc_batch_set_cycles(void *b, const uint64_t *cycles);
void orc_batch_set_counters(void *b, const uint64_t *in /* n x 4, as orc_batch_counters */);
void orc_batch_obs(void *b, uint8_t *out /* n x 128 */);
void *orc_batch_env(void *b, uint32_t i);
/* timed CPU baseline: returns wall seconds for frames x n env-frames. */
double orc_batch_bench(void *b, const uint8_t *actions, uint32_t frames, uint32_t nthreads);

/* ---- FlatMemory test bus (src/cpu.rs:69-90) with an access trace ---- */
typedef struct {
    uint8_t ram[65536];
    uint32_t trace[64]; /* (kind<<24)|(addr<<8)|value ; kind 0=read 1=write */
    uint32_t ntrace;
} orc_flat;
void *orc_flat_new(void);
void orc_flat_free(void *f);
void orc_flat_load(void *f, uint16_t base, const uint8_t *data, uint32_t len);
void orc_flat_poke(void *f, uint16_t addr, uint8_t v);
uint8_t orc_flat_peek(void *f, uint16_t addr);
/* regs: pc, sp, a, x, y, p (as uint32) */
void orc_flat_reset(void *f, uint32_t regs[6]);
int orc_flat_step(void *f, uint32_t regs[6], uint32_t *trace_out, uint32_t *ntrace);
/* Exhaustive ALU sweep: runs opcode `op` with immediate operand over all
 * (a, v, carry, decimal) = 256*256*2*2 inputs; out[i] = (p<<8)|a after. */
void orc_alu_sweep(uint8_t op, uint16_t *out);

/* ---- component hooks for tests (see orc.c) ---- */
void orc_component_run(const uint8_t *rom, uint32_t rom_len, uint8_t *state,
                       const uint32_t *ops, uint32_t nops, uint32_t *results);
void orc_pia_sweep(const uint16_t *in, const uint16_t *n, uint32_t k,
                   uint16_t *out_n, uint16_t *out_1, uint8_t *uf_seen);

/* ---- rendering console (orc_render.inc: src/tia.rs, src/atari.rs:107-318) ----
 * State: ORC_RSTATE_BYTES = the 222-byte AtariStateGpu layout (tia_clock in
 * TIA clocks 0..227), then visible_top @224, vblank_top @226, pending_top
 * @228 (u16), pending_frames @230, hmove_pending @231, inpt4_latch @232,
 * inpt5_latch @233, input_latch_enabled @234, zero to 255.  Frame: 160 x 192
 * palette indices. */
#define ORC_RSTATE_BYTES 256
#define ORC_FRAME_BYTES (160 * 192)
void *orc_r_new(const uint8_t *rom, uint32_t rom_len);        /* Atari::new */
void orc_r_free(void *h);
uint64_t orc_r_step(void *h, uint32_t action, uint32_t frames); /* set_action + frames x run_frame */
void orc_r_get_state(void *h, uint8_t *out);
void orc_r_set_state(void *h, const uint8_t *in);
void orc_r_get_frame(void *h, uint8_t *out);
void orc_r_set_frame(void *h, const uint8_t *in);
void orc_r_get_counters(void *h, uint64_t out[4]);
uint8_t orc_r_bus_read(void *h, uint16_t addr);
void orc_r_bus_write(void *h, uint16_t addr, uint8_t v);
void orc_r_tia_write(void *h, uint16_t addr, uint8_t v);
void orc_r_tia_ticks(void *h, uint32_t n);
void *orc_rbatch_new(const uint8_t *rom, uint32_t rom_len, uint32_t n);
void orc_rbatch_free(void *b);
void orc_rbatch_reset(void *b, uint32_t nthreads);            /* Atari::new + one run_frame */
void orc_rbatch_run(void *b, const uint8_t *actions, uint32_t frames, uint32_t nthreads);
void orc_rbatch_states(void *b, uint8_t *out);
void orc_rbatch_set_states(void *b, const uint8_t *in);
void orc_rbatch_frames(void *b, uint8_t *out);
void orc_rbatch_counters(void *b, uint64_t *out);
double orc_rbatch_bench(void *b, const uint8_t *actions, uint32_t frames, uint32_t nthreads);

#ifdef __cplusplus
}
#endif
#endif
