/* fake_plugin.c -- a plugin with the reference's ten exports and no GPU, for
 * CPU tests of the host's plugin manager (quasi_amd/host/manager.hpp).
 * Built by tests/test_plugin_manager.py with gcc in variants:
 *   FAKE_VERSION   the info version's major number and the readback value base;
 *   FAKE_ABI       what Q_plugin_abi_version returns (default 3);
 *   FAKE_NO_RENDER leaves Q_plugin_render out (a missing required symbol);
 *   FAKE_FAIL_CREATE makes Q_plugin_create return NULL.
 * create / destroy / render are reported through the context's log callback,
 * and the readback holds FAKE_VERSION * 1000 + frames rendered in every
 * channel, so a test can tell which build rendered and how often. */
#include <stdio.h>
#include <stdlib.h>

#include "quasi/plugin_interface.h"

#ifndef FAKE_VERSION
#define FAKE_VERSION 1
#endif
#ifndef FAKE_ABI
#define FAKE_ABI 3
#endif

struct Q_plugin_handle {
    Q_plugin_context* ctx;
    unsigned frames;
};

static void say(Q_plugin_context* ctx, const char* what, unsigned n) {
    char buf[96];
    snprintf(buf, sizeof buf, "fake v%d %s %u", FAKE_VERSION, what, n);
    if (ctx && ctx->log) ctx->log(ctx->host_data, buf);
}

uint32_t Q_plugin_abi_version(void) { return FAKE_ABI; }

Q_plugin_info Q_plugin_get_info(void) {
    Q_plugin_info i = {"fake", {FAKE_VERSION, 0, 0}, "GPU-free test plugin", "tests"};
    return i;
}

Q_plugin_handle* Q_plugin_create(Q_plugin_context* ctx) {
#ifdef FAKE_FAIL_CREATE
    say(ctx, "create refused", 0);
    return NULL;
#else
    Q_plugin_handle* h = (Q_plugin_handle*)calloc(1, sizeof *h);
    if (!h) return NULL;
    h->ctx = ctx;
    say(ctx, "create", 0);
    return h;
#endif
}

void Q_plugin_destroy(Q_plugin_handle* h) {
    if (!h) return;
    say(h->ctx, "destroy", h->frames);
    free(h);
}

void Q_plugin_update(Q_plugin_handle* h, float dt) {
    (void)h;
    (void)dt;
}

#ifndef FAKE_NO_RENDER
void Q_plugin_render(Q_plugin_handle* h, Q_render_frame* f) {
    (void)f;
    if (!h) return;
    h->frames++;
}
#endif

Q_readback_result Q_plugin_readback(Q_plugin_handle* h) {
    Q_readback_result r = {NULL, 0, 0, 0};
    if (!h || !h->ctx) return r;
    const uint32_t w = h->ctx->viewport_width, ht = h->ctx->viewport_height;
    r.data = (float*)malloc(sizeof(float) * 4u * w * ht);
    if (!r.data) return r;
    for (uint32_t i = 0; i < 4u * w * ht; ++i) r.data[i] = (float)(FAKE_VERSION * 1000 + h->frames);
    r.width = w;
    r.height = ht;
    r.channels = 4;
    return r;
}

void Q_plugin_readback_free(Q_readback_result* r) {
    if (!r) return;
    free(r->data);
    r->data = NULL;
}

Q_readback_aov_result Q_plugin_readback_aov(Q_plugin_handle* h) {
    Q_readback_aov_result r;
    for (int k = 0; k < Q_AOV_COUNT; ++k) {
        Q_readback_result b = Q_plugin_readback(h);
        r.buffers[k].data = b.data;
        r.buffers[k].width = b.width;
        r.buffers[k].height = b.height;
        r.buffers[k].channels = b.channels;
    }
    return r;
}

void Q_plugin_readback_aov_free(Q_readback_aov_result* r) {
    if (!r) return;
    for (int k = 0; k < Q_AOV_COUNT; ++k) {
        free(r->buffers[k].data);
        r->buffers[k].data = NULL;
    }
}
