// Case tables for Marching Cubes (MC) and Dual Marching Cubes (DMC), packed one 64-bit word
// per cube configuration. Shared by the HIP kernels and the CPU oracle; plain C, no device code.
//
// Provenance: the case contents (which edges each MC triangle uses, in which order and winding;
// which crossing edges form each DMC patch; which DMC configurations are ambiguous and which face
// neighbour resolves them) are transcribed from the reference tables
//   MC : firstMarchingCubesId / marchingCubesIds   src/cumc.cu:143-159
//   DMC: mcFirstPatchIndex / mcFirstEdgeIndex / mcEdgeIndex / dmcEdgeOffset / problematicConfigs
//        src/cudualmc.cu:121-501
// because identical topology and identical vertex/face order require identical case contents.
// The packing is this project's own; tests/test_tables.py expands it back and compares it against
// the reference text, and checks the geometric invariants independently.
//
// Corner numbering. MC (Bourke order, src/cumc.cu:107): corner i sits at
//   (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1) (1,0,1) (1,1,1) (0,1,1).
// DMC (Morton order, src/cudualmc.cu:95-104): corner i sits at (i&1, (i>>1)&1, (i>>2)&1).
// A corner bit is set when sdf >= iso.
//
// Edge numbering (both algorithms, src/cumc.cu:109-122): edge e starts at the cell-relative
// vertex DISO_EDGE_OFS[e] and runs one step along axis DISO_EDGE_AXIS[e] (0=x, 1=y, 2=z). An edge
// is therefore owned by the cell anchored at (cell + DISO_EDGE_OFS[e]) as that cell's
// axis-DISO_EDGE_AXIS[e] edge.
#ifndef DISO_TABLES_H
#define DISO_TABLES_H
#include <stdint.h>

#define DISO_EDGE_OFS_VALUES \
    {0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {0, 0, 0}, {0, 1, 0}, {1, 1, 0}, \
    {0, 1, 1}, {0, 1, 0}, {0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}
#define DISO_EDGE_AXIS_VALUES 0, 2, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1
static const int DISO_EDGE_OFS[12][3] = {DISO_EDGE_OFS_VALUES};
static const int DISO_EDGE_AXIS[12] = {DISO_EDGE_AXIS_VALUES};

// MC case word: bits [0,4) = triangle count (0..5); bits [4+4k, 8+4k) = k-th edge id of the flat
// triangle list (3 edges per triangle, k < 15).
#define DISO_MC_CASE_VALUES \
    0x0000000000000000ull, 0x0000000000008301ull, 0x0000000000001901ull, 0x0000000009318392ull, \
    0x0000000000005491ull, 0x0000000005498302ull, 0x0000000001405412ull, 0x0000003153548343ull, \
    0x0000000000004781ull, 0x0000000000734702ull, 0x0000000007840192ull, 0x0000007317194793ull, \
    0x0000000008579582ull, 0x0000005735309503ull, 0x0000001571780183ull, 0x0000000007355312ull, \
    0x0000000000003b21ull, 0x00000000028b0822ull, 0x000000000b231902ull, 0x00000098b9b21923ull, \
    0x000000000b235492ull, 0x000000549208b283ull, 0x000000b234015413ull, 0x00021b8b4b141544ull, \
    0x000000000478b232ull, 0x000000204247b273ull, 0x000000784b239013ull, 0x0009414717217b24ull, \
    0x00000023b5789583ull, 0x00027b2097295794ull, 0x0008075717013b24ull, 0x0000005711727b23ull, \
    0x0000000000002a11ull, 0x0000000008302a12ull, 0x0000000000a92a02ull, 0x000000a98a832a33ull, \
    0x0000000002a15492ull, 0x000000a125493083ull, 0x00000040242a54a3ull, 0x0003282a8a48a544ull, \
    0x0000000007842a12ull, 0x000000a127304703ull, 0x000000784a902a03ull, 0x000a327343a4a944ull, \
    0x0000002a18957853ull, 0x000573953930a124ull, 0x00080757a70a02a4ull, 0x00000073553a32a3ull, \
    0x0000000001ba3b12ull, 0x0000008ba8a10813ull, 0x000000ba9b903b03ull, 0x000000000ba88a92ull, \
    0x000000495ba13b13ull, 0x0008ba08a0a14954ull, 0x0004a5ba3a434034ull, 0x000000ba88a4a543ull, \
    0x00000047813ba1b3ull, 0x0007b40414b1ba14ull, 0x000ba93b93907844ull, 0x000000a9bb979473ull, \
    0x000785895ba33a14ull, 0x05750907b10a0ba5ull, 0x0bab030a58070575ull, 0x000000000a7ba572ull, \
    0x000000000000a651ull, 0x00000000065a8302ull, 0x00000000065a9012ull, 0x00000065a3198393ull, \
    0x000000000964a692ull, 0x000000308649a693ull, 0x000000640601a613ull, 0x00034864a43a31a4ull, \
    0x000000000a657842ull, 0x000000a650473073ull, 0x00000065a7841903ull, 0x00073147141965a4ull, \
    0x00000089a8a67863ull, 0x00067a9a0a707304ull, 0x0001a0a606806784ull, 0x0000003177161a63ull, \
    0x0000000003b265a2ull, 0x0000005a68b20823ull, 0x000000b2365a0193ull, 0x00021b8b9b19a654ull, \
    0x0000003b29a64963ull, 0x000b2820864aa494ull, 0x000640a60a01b234ull, 0x16461a14821b18b5ull, \
    0x000000a6523b4783ull, 0x0007b40424b265a4ull, 0x00065a7843b21904ull, 0x65a27b7214719415ull, \
    0x00067a9a8a78b234ull, 0x72027b70967a79a5ull, 0xb23867680a601a05ull, 0x00017b1b27161a64ull, \
    0x0000000005261252ull, 0x0000008305126523ull, 0x0000000260659053ull, 0x0005962636939834ull, \
    0x0000002642491293ull, 0x0002641241493084ull, 0x0000000006422402ull, 0x0000006422434833ull, \
    0x0000008472651253ull, 0x0006525127344304ull, 0x0005962606904784ull, 0x9737949325969265ull, \
    0x0008678916812614ull, 0x9262919670939735ull, 0x0000002600686783ull, 0x0000000003673262ull, \
    0x00000051353b65b3ull, 0x0005b6510b508b04ull, 0x0000393b9b59b654ull, 0x0000008b99b5b653ull, \
    0x000b631393696494ull, 0x18b8101b69141645ull, 0x00000040660b03b3ull, 0x0000000004b648b2ull, \
    0x000b631353657844ull, 0xb515b6b107b4b045ull, 0x7845b6b593b90395ull, 0x0009b6965b979474ull, \
    0x689867691b636135ull, 0x000000000b679102ull, 0x00006707860b03b4ull, 0x00000000000067b1ull, \
    0x00000000000076b1ull, 0x0000000006b73082ull, 0x0000000006b71902ull, 0x0000006b79831933ull, \
    0x000000000b764952ull, 0x000000b763085493ull, 0x000000b761540143ull, 0x000315835854b764ull, \
    0x000000000b4684b2ull, 0x00000004606b30b3ull, 0x000000190b846b43ull, 0x0006b33196394694ull, \
    0x000000b89b956b53ull, 0x000309b395b96b54ull, 0x000b561505b0b804ull, 0x00000015335b56b3ull, \
    0x0000000006372362ull, 0x0000006206087683ull, 0x0000009016237633ull, 0x0006879818616214ull, \
    0x0000009543762363ull, 0x0006207607085494ull, 0x0001541407623724ull, 0x8626878214858155ull, \
    0x0000004624238433ull, 0x0000000004624202ull, 0x0003826242840194ull, 0x0000006244292193ull, \
    0x0009566239638934ull, 0x0000002066050953ull, 0x8151808563828625ull, 0x0000000002562152ull, \
    0x00000000076ba122ull, 0x0000006b7a128303ull, 0x00000076b02a90a3ull, 0x000a982a82836b74ull, \
    0x00000076b4952a13ull, 0x000a12830549b764ull, 0x00040254252a76b4ull, 0x6b74a5a482a83285ull, \
    0x00000012a46b84b3ull, 0x000b364606302a14ull, 0x00002a0a96b84684ull, 0x3a9a32394b363465ull, \
    0x000b896b96952a14ull, 0x2a10b3b096b95695ull, 0x5b8b56580a525025ull, 0x00053252a35b56b4ull, \
    0x000000137176a163ull, 0x000a106a08607684ull, 0x00076aa907a03704ull, 0x00000098aa868763ull, \
    0x0006a73717a15494ull, 0x54916a6107608705ull, 0xa404a5a036a7a375ull, 0x0008a5854a868764ull, \
    0x00043846a34a13a4ull, 0x0000004600616a13ull, 0x34643836a0393a95ull, 0x0000000006946a92ull, \
    0x61316a6385696895ull, 0x00006a0a16050954ull, 0x00000000056a3802ull, 0x0000000000006a51ull, \
    0x0000000007ab5a72ull, 0x000000083ab75a73ull, 0x00000001975ab7a3ull, 0x000983931b75ab54ull, \
    0x0000009ab9b74973ull, 0x00074bab9b498304ull, 0x000b74401b41ab14ull, 0x43134841a74b4ab5ull, \
    0x000000ab8a845a43ull, 0x000a45ab34a30434ull, 0x000458b8a85a9014ull, 0x4aba454b39414315ull, \
    0x000000000ab8a892ull, 0x000000ab99b0b303ull, 0x000000b8aa818013ull, 0x000000000b1ab312ull, \
    0x00000037535a23a3ull, 0x00008775a07a20a4ull, 0x000a257535231904ull, 0x298921287a252755ull, \
    0x0003a2374a349a44ull, 0x79a9747a28707205ull, 0xa373a2a741a0a405ull, 0x000000000874a212ull, \
    0x000238a284a85a44ull, 0x00000004224a45a3ull, 0x1903a2a385a84585ull, 0x00042141924a45a4ull, \
    0x0000009a88a3a233ull, 0x000000000a09a202ull, 0x000a80a018a3a234ull, 0x000000000000a211ull, \
    0x000000751712b723ull, 0x000751b71b128304ull, 0x00072b0292797594ull, 0x27572b2593282985ull, \
    0x000491741271b724ull, 0x830974791b712b15ull, 0x0000000244272b73ull, 0x0002482834272b74ull, \
    0x00012bb841b45144ull, 0xb040b3b452b1b515ull, 0x50205952b4585b85ull, 0x0000000002b34592ull, \
    0x00000089bb929123ull, 0x0009b3930b929124ull, 0x00000000082b8022ull, 0x000000000000b321ull, \
    0x0000000003753512ull, 0x0000005177181083ull, 0x0000007533505903ull, 0x0000000005875982ull, \
    0x0000003711797493ull, 0x0001741497181084ull, 0x0000000007037402ull, 0x0000000000007481ull, \
    0x0000001355343843ull, 0x0000000004104512ull, 0x0003593905343844ull, 0x0000000000004591ull, \
    0x0000000003913892ull, 0x0000000000009101ull, 0x0000000000003801ull, 0x0000000000000000ull,
static const uint64_t DISO_MC_CASE[256] = {DISO_MC_CASE_VALUES};

// DMC case word: bits [3e, 3e+3) = patch index (0..3) that crossing edge e belongs to, 7 when
// edge e does not cross; bits [36,39) = patch count (0..4); bits [39,42) = ambiguity direction
// d (0..5: face neighbour along axis d>>1, towards + when d&1) or 7 when the case is not
// ambiguous; bits [42+3k, 45+3k) = number of edges in patch k (derived from the edge fields).
// Within a patch the edges are taken in ascending edge id.
#define DISO_DMC_CASE_VALUES \
    0x0000038fffffffffull, 0x00000f9ff8fff1f8ull, 0x00000f9fc7ffffc0ull, 0x0000139fc0fff1c7ull, \
    0x00000f9ff81f8fffull, 0x0000139fff1f81f8ull, 0x00006fafc13f9fc0ull, 0x0000179fc71f81c7ull, \
    0x00000f9fc7fc0fffull, 0x00006fafc8fc91f8ull, 0x0000139ffffc0fc0ull, 0x0000179ff8fc01c7ull, \
    0x0000139fc01c7fffull, 0x0000179fc71c71f8ull, 0x0000179ff81c7fc0ull, 0x0000139fff1c71c7ull, \
    0x00000f91fffff03full, 0x00001391f8fffe38ull, 0x00006fa3c7fff240ull, 0x00001791c0fffe07ull, \
    0x00006fa3f81f827full, 0x00001791ff1f8e38ull, 0x00036fb5c13f9480ull, 0x00001b91c71f8e07ull, \
    0x00006fa3c7fc027full, 0x000073a1c8fc9e38ull, 0x000073a3fffc0240ull, 0x00001b91f8fc0e07ull, \
    0x000073a3c01c727full, 0x00001b91c71c7e38ull, 0x000077a3f81c7240ull, 0x00001791ff1c7e07ull, \
    0x00000f9e3ffffe07ull, 0x00006fae78fff048ull, 0x0000139e07fffe38ull, 0x0000179e00fff03full, \
    0x00006fae781f8e4full, 0x000073ae7f1f8048ull, 0x000073ae013f9e38ull, 0x00001b9e071f803full, \
    0x00006fae47fc0e4full, 0x00036fbe88fc9090ull, 0x0000179e3ffc0e38ull, 0x00001b9e38fc003full, \
    0x000073ae401c7e4full, 0x000077ae471c7048ull, 0x00001b9e381c7e38ull, 0x0000179e3f1c703full, \
    0x000013903ffff1c7ull, 0x0000179038ffffc0ull, 0x0000179007fff1f8ull, 0x0000139000ffffffull, \
    0x00008fa2781f83cfull, 0x00001b903f1f8fc0ull, 0x000077a0013f91f8ull, 0x00001790071f8fffull, \
    0x00008fa247fc03cfull, 0x000077a008fc9fc0ull, 0x00001b903ffc01f8ull, 0x0000179038fc0fffull, \
    0x000093a2401c73cfull, 0x00001c90071c7fc0ull, 0x00001c10381c71f8ull, 0x000013903f1c7fffull, \
    0x00000f91ff03ffffull, 0x00006fa3f827f1f8ull, 0x00006fa3c727ffc0ull, 0x000073a3c027f1c7ull, \
    0x00001391f8e38fffull, 0x00001791ffe381f8ull, 0x00008fa3c1e79fc0ull, 0x00001b91c7e381c7ull, \
    0x00006fa3c7240fffull, 0x00036fb5c84891f8ull, 0x000073a3ff240fc0ull, 0x000077a3f82401c7ull, \
    0x00001791c0e07fffull, 0x00001b91c7e071f8ull, 0x00001b91f8e07fc0ull, 0x00001791ffe071c7ull, \
    0x0000139fff03f03full, 0x0000179ff803fe38ull, 0x00008fafc727f240ull, 0x00001b9fc003fe07ull, \
    0x0000179ff8e3803full, 0x0000139fffe38e38ull, 0x0000afafc1e79240ull, 0x0000179fc7e38e07ull, \
    0x00008fafc724027full, 0x000077afc8009e38ull, 0x000093afff240240ull, 0x00001d9ff8000e07ull, \
    0x00001b9fc0e0703full, 0x0000179fc7e07e38ull, 0x00001d1ff8e07000ull, 0x0000139fffe07e07ull, \
    0x00006fa23f27fe07ull, 0x00036fb4784bf048ull, 0x000073a20727fe38ull, 0x000077a20027f03full, \
    0x000073a078e38e4full, 0x000077a07fe38048ull, 0x000093a201e79e38ull, 0x00001e9007e3803full, \
    0x00036fb447480e4full, 0x001b6fc6886c9090ull, 0x000077a23f240e38ull, 0x00007ba23824003full, \
    0x000077a040e07e4full, 0x00007ba047e07048ull, 0x00001e9038e07e38ull, 0x00001a903fe0703full, \
    0x0000179e3f03f1c7ull, 0x00001b9e3803ffc0ull, 0x00001b9e0703f1f8ull, 0x0000179e0003ffffull, \
    0x00001b9e38e381c7ull, 0x0000179e3fe38fc0ull, 0x00001e1e00e381f8ull, 0x0000139e07e38fffull, \
    0x0000afae472403cfull, 0x00007bae08009fc0ull, 0x00001d9e3f0001f8ull, 0x0000199e38000fffull, \
    0x00001c9e00e071c7ull, 0x0000189e07e07fc0ull, 0x00006fae78e4f1f8ull, 0x00000f9e3fe07fffull, \
    0x00000f9e3fe07fffull, 0x00006fae78e4f1f8ull, 0x00006fae47e4ffc0ull, 0x000073ae40e4f1c7ull, \
    0x00006fae78048fffull, 0x000073ae7f0481f8ull, 0x00036fbe81291fc0ull, 0x000077ae470481c7ull, \
    0x0000139e07e38fffull, 0x00008fae48e791f8ull, 0x0000179e3fe38fc0ull, 0x00001b9e38e381c7ull, \
    0x0000179e0003ffffull, 0x00001b9e0703f1f8ull, 0x00001b9e3803ffc0ull, 0x0000179e3f03f1c7ull, \
    0x00006fa07fe4f03full, 0x000073a078e4fe38ull, 0x00036fb287e97240ull, 0x000077a040e4fe07ull, \
    0x00036fb2b809027full, 0x000077a07f048e38ull, 0x001b6fc4c12d9480ull, 0x00007ba047048e07ull, \
    0x000073a207e3827full, 0x000093a048e79e38ull, 0x000077a23fe38240ull, 0x00001e9038e38e07ull, \
    0x000077a20003f27full, 0x00001e900703fe38ull, 0x00007ba23803f240ull, 0x00001a903f03fe07ull, \
    0x0000139fffe07e07ull, 0x00008faff8e4f048ull, 0x0000179fc7e07e38ull, 0x00001b9fc0e0703full, \
    0x00008faff8048e4full, 0x000093afff048048ull, 0x000077afc1201e38ull, 0x00001d9fc700003full, \
    0x0000179fc7e38e07ull, 0x0000afafc8e79048ull, 0x0000139fffe38e38ull, 0x0000179ff8e3803full, \
    0x00001b9fc003fe07ull, 0x00001d1fc703f000ull, 0x0000179ff803fe38ull, 0x0000139fff03f03full, \
    0x00001791ffe071c7ull, 0x00001b91f8e07fc0ull, 0x00001b91c7e071f8ull, 0x00001791c0e07fffull, \
    0x0000afa3f80483cfull, 0x00001d91ff000fc0ull, 0x00007ba1c12011f8ull, 0x00001991c7000fffull, \
    0x00001b91c7e381c7ull, 0x00001e11c0e38fc0ull, 0x00001791ffe381f8ull, 0x00001391f8e38fffull, \
    0x00001c11c003f1c7ull, 0x00006fa3c727ffc0ull, 0x00001811f803f1f8ull, 0x00000f91ff03ffffull, \
    0x000013903f1c7fffull, 0x00008fa2783cf1f8ull, 0x00008fa2473cffc0ull, 0x000093a2403cf1c7ull, \
    0x0000179038fc0fffull, 0x00001b903ffc01f8ull, 0x0000afa241fc9fc0ull, 0x00001c9007fc01c7ull, \
    0x00001790071f8fffull, 0x0000afa2483f91f8ull, 0x00001b903f1f8fc0ull, 0x00001c10381f81c7ull, \
    0x0000139000ffffffull, 0x0000179007fff1f8ull, 0x0000179038ffffc0ull, 0x000013903ffff1c7ull, \
    0x0000179e3f1c703full, 0x00001b9e381c7e38ull, 0x0000afae473cf240ull, 0x00001c9e001c7e07ull, \
    0x00001b9e38fc003full, 0x0000179e3ffc0e38ull, 0x0000cfae41fc9240ull, 0x0000189e07fc0e07ull, \
    0x00001b9e071f803full, 0x00001e1e001f8e38ull, 0x00001d1e3f1f8000ull, 0x00006fae781f8e4full, \
    0x0000179e00fff03full, 0x0000139e07fffe38ull, 0x0000191e38fff000ull, 0x00000f9e3ffffe07ull, \
    0x00001791ff1c7e07ull, 0x0000afa3f83cf048ull, 0x00001b91c71c7e38ull, 0x00001c11c01c703full, \
    0x00001b91f8fc0e07ull, 0x00001d11fffc0000ull, 0x00001e11c0fc0e38ull, 0x00006fa3c7fc027full, \
    0x00001b91c71f8e07ull, 0x0000cfa3c83f9048ull, 0x00001791ff1f8e38ull, 0x00001811f81f803full, \
    0x00001791c0fffe07ull, 0x00001911c7fff000ull, 0x00001391f8fffe38ull, 0x00000f91fffff03full, \
    0x0000139fff1c71c7ull, 0x0000179ff81c7fc0ull, 0x0000179fc71c71f8ull, 0x0000139fc01c7fffull, \
    0x0000179ff8fc01c7ull, 0x0000139ffffc0fc0ull, 0x00001a1fc0fc01f8ull, 0x00000f9fc7fc0fffull, \
    0x0000179fc71f81c7ull, 0x00001a1fc01f8fc0ull, 0x0000139fff1f81f8ull, 0x00000f9ff81f8fffull, \
    0x0000139fc0fff1c7ull, 0x00000f9fc7ffffc0ull, 0x00000f9ff8fff1f8ull, 0x0000038fffffffffull,
static const uint64_t DISO_DMC_CASE[256] = {DISO_DMC_CASE_VALUES};

// DMC quad stencil (src/cudualmc.cu:504-513): for a crossing edge owned by a cell and of type
// ty = axis + 3*(edge leaves the inside, i.e. sdf(start) >= iso > sdf(end)), the quad's corners
// k = 0..3 are the DMC vertices of cells (cell + DISO_QUAD[ty][k][0..2]) at their edge
// DISO_QUAD[ty][k][3].
#define DISO_QUAD_VALUES \
    {{0, 0, 0, 0}, {0, -1, 0, 4}, {0, -1, -1, 6}, {0, 0, -1, 2}}, \
    {{0, 0, 0, 8}, {0, 0, -1, 11}, {-1, 0, -1, 10}, {-1, 0, 0, 9}}, \
    {{0, 0, 0, 3}, {-1, 0, 0, 1}, {-1, -1, 0, 5}, {0, -1, 0, 7}}, \
    {{0, 0, 0, 0}, {0, 0, -1, 2}, {0, -1, -1, 6}, {0, -1, 0, 4}}, \
    {{0, 0, 0, 8}, {-1, 0, 0, 9}, {-1, 0, -1, 10}, {0, 0, -1, 11}}, \
    {{0, 0, 0, 3}, {0, -1, 0, 7}, {-1, -1, 0, 5}, {-1, 0, 0, 1}}
static const int DISO_QUAD[6][4][4] = {DISO_QUAD_VALUES};

#define DISO_MC_NTRI(w) ((int)((w) & 0xFull))
#define DISO_MC_EDGE(w, k) ((int)(((w) >> (4 + 4 * (k))) & 0xFull))
#define DISO_DMC_PATCH_OF(w, e) ((int)(((w) >> (3 * (e))) & 0x7ull))
#define DISO_DMC_NPATCH(w) ((int)(((w) >> 36) & 0x7ull))
#define DISO_DMC_DIR(w) ((int)(((w) >> 39) & 0x7ull))
#define DISO_DMC_PATCH_SIZE(w, k) ((int)(((w) >> (42 + 3 * (k))) & 0x7ull))

#endif
