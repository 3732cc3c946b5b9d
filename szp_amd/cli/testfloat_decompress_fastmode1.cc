// Decompresses <file>.szp with szp_decompress() (libszp_amd.so), compares the
// values with the original <file>, reports the error metrics and writes
// <file>.szp.out.  Arguments and printed metrics follow the reference's
// example/testfloat_decompress_fastmode1.cc:
//   testfloat_decompress_fastmode1 srcFilePath.szp nbEle blockSize {float|double}
//                                  {random|nonrandom} [repetitions] [warmup_runs] [output_csv]
#include "cli_common.h"

struct Quality {
    double max_abs, max_rel, max_pw_rel, psnr, nrmse, ac_eff, ratio;
};

// Error metrics of the decompressed values d against the original values o.
template <typename V>
static Quality evaluate(const V *o, const V *d, size_t n, size_t byteLength)
{
    Quality q{0, 0, 0, 0, 0, 0, 0};
    double mn = n ? (double)o[0] : 0, mx = mn;
    double s1 = 0, s2 = 0;
    for (size_t i = 0; i < n; ++i) {
        s1 += o[i];
        s2 += d[i];
    }
    const double m1 = n ? s1 / n : 0, m2 = n ? s2 / n : 0;
    double v1 = 0, v2 = 0, cov = 0, sq = 0;
    for (size_t i = 0; i < n; ++i) {
        const double a = o[i], b = d[i];
        mn = a < mn ? a : mn;
        mx = a > mx ? a : mx;
        const double e = fabs(b - a);
        if (a != 0 && e / fabs(a) > q.max_pw_rel) q.max_pw_rel = e / fabs(a);
        if (e > q.max_abs) q.max_abs = e;
        cov += (a - m1) * (b - m2);
        v1 += (a - m1) * (a - m1);
        v2 += (b - m2) * (b - m2);
        sq += e * e;
    }
    const double sd1 = n ? sqrt(v1 / n) : 0, sd2 = n ? sqrt(v2 / n) : 0;
    const double mse = n ? sq / n : 0, range = mx - mn;
    q.ac_eff = sd1 * sd2 == 0 ? 0 : (n ? cov / n : 0) / (sd1 * sd2);
    q.max_rel = range == 0 ? 0 : q.max_abs / range;
    q.psnr = mse == 0 ? 999 : 20 * log10(range) - 10 * log10(mse);
    q.nrmse = range == 0 ? 0 : sqrt(mse) / range;
    q.ratio = byteLength ? (double)n * sizeof(V) / byteLength : 0;
    printf("Min=%.20G, Max=%.20G, range=%.20G\n", mn, mx, range);
    printf("Max absolute error = %.10f\n", q.max_abs);
    printf("Max relative error = %f\n", q.max_rel);
    printf("Max pw relative error = %f\n", q.max_pw_rel);
    printf("PSNR = %f, NRMSE= %.20G\n", q.psnr, q.nrmse);
    printf("acEff=%f\n", q.ac_eff);
    printf("compressionRatio = %f\n", q.ratio);
    return q;
}

static void usage()
{
    printf("Usage: testfloat_decompress_fastmode1 [srcFilePath] [nbEle] [block size] [data_type] "
           "[access_type] [repetitions] [warmup_runs] [output_csv]\n");
    printf("Example: testfloat_decompress_fastmode1 field.f32.szp 8192 128 float nonrandom 10 2 results.csv\n");
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        usage();
        return 0;
    }
    const char *zip = argv[1];
    size_t nbEle = strtoull(argv[2], nullptr, 10);
    const int blockSize = atoi(argv[3]);
    const int dtype = cli::parse_data_type(argv[4]);
    const int access = cli::parse_access(argv[5]);
    const int reps = argc >= 7 ? atoi(argv[6]) : 1;
    const int warmup = argc >= 8 ? atoi(argv[7]) : 0;
    const char *csv = argc >= 9 ? argv[8] : nullptr;
    if (dtype < 0) {
        printf("Error: data type must be 'float' or 'double'!\n");
        return 0;
    }
    if (access < 0) {
        printf("Error: access type must be 'random' or 'nonrandom'!\n");
        return 0;
    }
    const size_t zl = strlen(zip);
    if (zl < 5 || strcmp(zip + zl - 4, ".szp") != 0) {
        printf("Error: %s does not end in .szp\n", zip);
        return 0;
    }
    printf("SZp Decompression Benchmark Configuration\n");
    printf("=========================================\n");
    printf("Compressed file: %s\nElements: %zu\nBlock size: %d\nData type: %s\nAccess type: %s\n",
           zip, nbEle, blockSize, argv[4], argv[5]);
    printf("Repetitions: %d\nWarmup runs: %d\n", reps, warmup);
    if (csv) printf("CSV output: %s\n", csv);
    printf("=========================================\n\n");

    std::vector<char> out_path(zl + 8), ori_path(zl + 1);
    snprintf(out_path.data(), out_path.size(), "%s.out", zip);
    memcpy(ori_path.data(), zip, zl - 4);
    ori_path[zl - 4] = '\0';

    int status = SZ_SCES;
    size_t byteLength = 0;
    printf("Reading compressed file...\n");
    unsigned char *bytes = szp_readByteData((char *)zip, &byteLength, &status);
    if (status != SZ_SCES) {
        printf("Error: %s cannot be read!\n", zip);
        return 0;
    }
    printf("Reading original file for comparison...\n");
    size_t total = 0;
    void *ori = dtype == SZ_FLOAT ? (void *)szp_readFloatData(ori_path.data(), &total, &status)
                                  : (void *)szp_readDoubleData(ori_path.data(), &total, &status);
    if (status != SZ_SCES) {
        printf("Error: %s cannot be read!\n", ori_path.data());
        free(bytes);
        return 0;
    }
    if (nbEle > total) {
        printf("Warning: Specified element count (%zu) exceeds file size (%zu). Using file size.\n",
               nbEle, total);
        nbEle = total;
    }

    if (warmup > 0) printf("\nPerforming %d warmup runs...\n", warmup);
    for (int i = 0; i < warmup; ++i) {
        const double t0 = cli::now_seconds();
        void *data = szp_decompress(access, dtype, bytes, byteLength, nbEle, blockSize);
        printf("  Warmup run %d of %d: time=%f\n", i + 1, warmup, cli::now_seconds() - t0);
        free(data);
    }
    std::vector<double> times, psnrs;
    std::vector<Quality> qs;
    printf("\nStarting %d benchmark runs...\n", reps);
    for (int i = 0; i < reps; ++i) {
        printf("\nRun %d of %d:\n", i + 1, reps);
        const double t0 = cli::now_seconds();
        void *data = szp_decompress(access, dtype, bytes, byteLength, nbEle, blockSize);
        const double t = cli::now_seconds() - t0;
        if (!data) {
            printf("Error: decompression failed\n");
            free(bytes);
            free(ori);
            return 1;
        }
        printf("timecost=%f\n", t);
        const Quality q = dtype == SZ_FLOAT
                              ? evaluate((const float *)ori, (const float *)data, nbEle, byteLength)
                              : evaluate((const double *)ori, (const double *)data, nbEle, byteLength);
        times.push_back(t);
        psnrs.push_back(q.psnr);
        qs.push_back(q);
        if (i == reps - 1) {
            printf("  Writing decompressed data to %s\n", out_path.data());
            if (dtype == SZ_FLOAT)
                szp_writeFloatData_inBytes((float *)data, nbEle, out_path.data(), &status);
            else
                szp_writeDoubleData_inBytes((double *)data, nbEle, out_path.data(), &status);
            if (status != SZ_SCES) printf("Error: %s cannot be written!\n", out_path.data());
        }
        free(data);
    }
    if (times.size() > 1) {
        printf("\n======= Performance Statistics (excluding warmup runs) =======\n");
        cli::print_stat("Time (seconds)", times);
        cli::print_stat("PSNR", psnrs);
        printf("===========================================================\n");
    }
    if (csv) {
        FILE *f = fopen(csv, "w");
        if (!f) {
            printf("Error: Cannot open file %s for writing\n", csv);
        } else {
            fprintf(f, "Run,TimeSeconds,MaxAbsErr,MaxRelErr,MaxPwRelErr,PSNR,NRMSE,AccEff,CompressionRatio\n");
            for (size_t i = 0; i < qs.size(); ++i)
                fprintf(f, "%zu,%f,%f,%f,%f,%f,%e,%f,%f\n", i + 1, times[i], qs[i].max_abs,
                        qs[i].max_rel, qs[i].max_pw_rel, qs[i].psnr, qs[i].nrmse, qs[i].ac_eff,
                        qs[i].ratio);
            fclose(f);
            printf("Results written to %s\n", csv);
        }
    }
    free(bytes);
    free(ori);
    return 0;
}
